// CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h). Built with -ffp-contract=off so that
// the only fused multiply-adds are the explicit fmaf() calls that mirror the reference's FMA use.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <queue>
#include <random>
#include <vector>

// The reference's 4-lane tail (DistanceFunctions.hpp:325-339) is written as add(mul(..)). With
// -mfma and GCC's default -ffp-contract=fast the compiler fuses it; the reference module built
// by oracle/Makefile therefore computes a fused tail, and so does this restatement.
static const bool kTailFused = true;

extern "C" float or_dist_l2(const float *x, const float *y, int64_t d) {
    // DistanceFunctions.hpp:304-345: 8 FMA accumulators over full blocks of 8 ...
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t i = 0;
    for (; d - i >= 8; i += 8)
        for (int j = 0; j < 8; ++j) {
            float t = x[i + j] - y[i + j];
            acc[j] = std::fmaf(t, t, acc[j]);
        }
    // ... folded high half onto low half (:321-322) ...
    float m[4];
    for (int j = 0; j < 4; ++j) m[j] = acc[4 + j] + acc[j];
    // ... a 4-wide block (:324-332) and a masked 1..3 tail (:334-339) ...
    for (int pass = 0; pass < 2; ++pass) {
        int64_t rem = d - i;
        int w = pass == 0 ? (rem >= 4 ? 4 : 0) : (int)rem;
        if (w == 0) continue;
        for (int j = 0; j < 4; ++j) {
            float t = j < w ? x[i + j] - y[i + j] : 0.0f;
            m[j] = kTailFused ? std::fmaf(t, t, m[j]) : m[j] + t * t;
        }
        i += w;
    }
    // ... two horizontal adds (:341-342) and a correctly rounded sqrt.
    float r = (m[0] + m[1]) + (m[2] + m[3]);
    return std::sqrt(r);
}

extern "C" float or_dist_l1(const float *a, const float *b, int64_t d) {
    // DistanceFunctions.hpp:347-382: 8 lane sums of |a-b|, summed lane 0..7, then a scalar tail.
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t i = 0;
    for (; i + 8 <= d; i += 8)
        for (int j = 0; j < 8; ++j) s[j] = s[j] + std::fabs(a[i + j] - b[i + j]);
    float total = 0.0f;
    for (int j = 0; j < 8; ++j) total += s[j];
    for (; i < d; ++i) total += std::fabs(a[i] - b[i]);
    return total;
}

extern "C" float or_dist_linf(const float *a, const float *b, int64_t d) {
    // DistanceFunctions.hpp:384-420: max |a-b|; max is exact so lane order does not matter.
    float m = 0.0f;
    for (int64_t i = 0; i < d; ++i) m = std::max(m, std::fabs(a[i] - b[i]));
    return m;
}

extern "C" int64_t or_dist_hamming(const uint8_t *a, const uint8_t *b, int64_t nbytes) {
    int64_t h = 0;
    for (int64_t i = 0; i < nbytes; ++i) h += __builtin_popcount((unsigned)(a[i] ^ b[i]));
    return h;
}

static float dist_by_metric(int metric, const float *a, const float *b, int64_t d) {
    switch (metric) {
    case OR_L1: return or_dist_l1(a, b, d);
    case OR_LINF: return or_dist_linf(a, b, d);
    default: return or_dist_l2(a, b, d);
    }
}

template <typename D>
static void topk_lex(std::vector<std::pair<D, int64_t>> &all, int64_t k, int64_t *oi, D *od) {
    int64_t kk = std::min<int64_t>(k, (int64_t)all.size());
    std::partial_sort(all.begin(), all.begin() + kk, all.end());
    for (int64_t j = 0; j < kk; ++j) {
        od[j] = all[j].first;
        oi[j] = all[j].second;
    }
}

extern "C" void or_knn_float(int metric, const float *X, int64_t n, int64_t d, const float *Q, int64_t nq,
                             int64_t k, int64_t *out_idx, float *out_dist) {
    int64_t kk = std::min(k, n);
#pragma omp parallel for schedule(dynamic)
    for (int64_t q = 0; q < nq; ++q) {
        std::vector<std::pair<float, int64_t>> all(n);
        for (int64_t i = 0; i < n; ++i) all[i] = {dist_by_metric(metric, Q + q * d, X + i * d, d), i};
        topk_lex(all, k, out_idx + q * kk, out_dist + q * kk);
    }
}

extern "C" void or_knn_hamming(const uint8_t *X, int64_t n, int64_t nbytes, const uint8_t *Q, int64_t nq,
                               int64_t k, int64_t *out_idx, int64_t *out_dist) {
    int64_t kk = std::min(k, n);
#pragma omp parallel for schedule(dynamic)
    for (int64_t q = 0; q < nq; ++q) {
        std::vector<std::pair<int64_t, int64_t>> all(n);
        for (int64_t i = 0; i < n; ++i) all[i] = {or_dist_hamming(Q + q * nbytes, X + i * nbytes, nbytes), i};
        topk_lex(all, k, out_idx + q * kk, out_dist + q * kk);
    }
}

extern "C" int32_t or_ivfbin_build(const uint8_t *X, int64_t n, int64_t nbytes, int32_t nlist, int32_t max_iter,
                                   uint32_t seed, uint8_t *centroids, int32_t *labels_out) {
    // BinaryIVF.hpp:127-227. The random draws (mt19937 + libstdc++ distributions) are made in the
    // same order as the reference: pick, then wsel when total > 0, per k-means++ step; one pick
    // per empty cluster in cluster order per Lloyd iteration.
    const int32_t k = (int32_t)std::min<int64_t>(nlist, n);
    const uint8_t *row = X;
    auto R = [&](int64_t i) { return row + i * nbytes; };
    std::mt19937 rng(seed);
    std::uniform_int_distribution<size_t> pick(0, (size_t)n - 1);
    std::vector<uint8_t> cent((size_t)k * nbytes);
    auto C = [&](int64_t c) { return cent.data() + c * nbytes; };
    std::memcpy(C(0), R((int64_t)pick(rng)), nbytes);
    std::vector<int64_t> min_d(n, std::numeric_limits<int64_t>::max());
    for (int32_t ci = 1; ci < k; ++ci) {
        const uint8_t *prev = C(ci - 1);
        for (int64_t i = 0; i < n; ++i) min_d[i] = std::min(min_d[i], or_dist_hamming(R(i), prev, nbytes));
        int64_t total = 0;
        for (int64_t v : min_d) total += v;
        size_t chosen = pick(rng);
        if (total > 0) {
            std::uniform_int_distribution<int64_t> wsel(0, total - 1);
            int64_t r = wsel(rng), cum = 0;
            for (int64_t i = 0; i < n; ++i) {
                cum += min_d[i];
                if (cum > r) { chosen = (size_t)i; break; }
            }
        }
        std::memcpy(C(ci), R((int64_t)chosen), nbytes);
    }
    std::vector<int32_t> labels(n, 0);
    std::vector<int32_t> bits((size_t)k * nbytes * 8), counts(k);
    for (int32_t it = 0; it < max_iter; ++it) {
        int64_t changed = 0;
        for (int64_t i = 0; i < n; ++i) {
            int64_t bd = std::numeric_limits<int64_t>::max();
            int32_t bc = 0;
            for (int32_t c = 0; c < k; ++c) {
                int64_t dd = or_dist_hamming(R(i), C(c), nbytes);
                if (dd < bd) { bd = dd; bc = c; }
            }
            if (labels[i] != bc) { labels[i] = bc; ++changed; }
        }
        if (changed == 0) break;
        std::fill(bits.begin(), bits.end(), 0);
        std::fill(counts.begin(), counts.end(), 0);
        for (int64_t i = 0; i < n; ++i) {
            int32_t c = labels[i];
            ++counts[c];
            for (int64_t b = 0; b < nbytes; ++b)
                for (int t = 0; t < 8; ++t)
                    if ((R(i)[b] >> t) & 1) ++bits[((size_t)c * nbytes + b) * 8 + t];
        }
        for (int32_t c = 0; c < k; ++c) {
            if (counts[c] == 0) {
                std::memcpy(C(c), R((int64_t)pick(rng)), nbytes);
                continue;
            }
            int32_t half = counts[c] / 2;
            for (int64_t b = 0; b < nbytes; ++b) {
                uint8_t v = 0;
                for (int t = 0; t < 8; ++t)
                    if (bits[((size_t)c * nbytes + b) * 8 + t] > half) v |= (uint8_t)(1u << t);
                C(c)[b] = v;
            }
        }
    }
    std::memcpy(centroids, cent.data(), cent.size());
    std::memcpy(labels_out, labels.data(), n * sizeof(int32_t));
    return k;
}

extern "C" void or_ivfbin_search(const uint8_t *cent, int32_t nc, const int64_t *list_off, const int32_t *list_ids,
                                 const uint8_t *X, int64_t nbytes, const uint8_t *Q, int64_t nq, int32_t nprobe,
                                 int64_t k, int64_t *out_idx, int64_t *out_dist, int32_t *out_count) {
    // BinaryIVF.hpp:65-111, literally: (dist, cid) partial sort of the centroids, then a max-heap
    // of (dist, id) with strict admission over the probed lists, emitted ascending.
    int32_t np = std::min(nprobe, nc);
    for (int64_t q = 0; q < nq; ++q) {
        const uint8_t *qq = Q + q * nbytes;
        std::vector<std::pair<int64_t, int32_t>> cd(nc);
        for (int32_t c = 0; c < nc; ++c) cd[c] = {or_dist_hamming(qq, cent + (int64_t)c * nbytes, nbytes), c};
        std::partial_sort(cd.begin(), cd.begin() + np, cd.end());
        std::priority_queue<std::pair<int64_t, int64_t>> heap;
        for (int32_t p = 0; p < np; ++p) {
            int32_t c = cd[p].second;
            for (int64_t j = list_off[c]; j < list_off[c + 1]; ++j) {
                int64_t id = list_ids[j];
                int64_t dd = or_dist_hamming(qq, X + id * nbytes, nbytes);
                if ((int64_t)heap.size() < k || dd < heap.top().first) {
                    heap.push({dd, id});
                    if ((int64_t)heap.size() > k) heap.pop();
                }
            }
        }
        int32_t cnt = (int32_t)heap.size();
        out_count[q] = cnt;
        for (int64_t j = 0; j < k; ++j) { out_idx[q * k + j] = -1; out_dist[q * k + j] = -1; }
        for (int32_t j = cnt - 1; j >= 0; --j) {
            out_idx[q * k + j] = heap.top().second;
            out_dist[q * k + j] = heap.top().first;
            heap.pop();
        }
    }
}

extern "C" void or_kmeans_l2(const float *X, int64_t n, int64_t d, int64_t k, int64_t max_iter, uint32_t seed,
                             int32_t *labels_out, float *centroids_out) {
    // KMeans.hpp:25-126: k-means++ (pick, then one unif01 draw per new centroid), Lloyd with
    // strict-< argmin over sqrt distances, serial per-point centroid sums scaled by 1/count,
    // empty clusters reseeded with one pick each in cluster order.
    std::mt19937 rng(seed);
    std::uniform_int_distribution<size_t> pick(0, (size_t)n - 1);
    std::uniform_real_distribution<float> unif01(0.f, 1.f);
    std::vector<float> cent((size_t)k * d);
    std::memcpy(cent.data(), X + (int64_t)pick(rng) * d, d * sizeof(float));
    std::vector<float> min_sq(n, std::numeric_limits<float>::max());
    for (int64_t ci = 1; ci < k; ++ci) {
        const float *prev = cent.data() + (ci - 1) * d;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            float v = or_dist_l2(X + i * d, prev, d);
            v = v * v;
            if (v < min_sq[i]) min_sq[i] = v;
        }
        double total = 0.0;
        for (float v : min_sq) total += (double)v;
        double r = (double)unif01(rng) * total, cum = 0.0;
        int64_t chosen = n - 1;
        for (int64_t i = 0; i < n; ++i) {
            cum += (double)min_sq[i];
            if (cum >= r) { chosen = i; break; }
        }
        std::memcpy(cent.data() + ci * d, X + chosen * d, d * sizeof(float));
    }
    std::vector<int32_t> labels(n, 0);
    std::vector<float> nc((size_t)k * d);
    std::vector<int64_t> counts(k);
    for (int64_t it = 0; it < max_iter; ++it) {
        int64_t changed = 0;
#pragma omp parallel for schedule(static) reduction(+ : changed)
        for (int64_t i = 0; i < n; ++i) {
            float bd = std::numeric_limits<float>::max();
            int32_t bc = 0;
            for (int64_t c = 0; c < k; ++c) {
                float dd = or_dist_l2(X + i * d, cent.data() + c * d, d);
                if (dd < bd) { bd = dd; bc = (int32_t)c; }
            }
            if (labels[i] != bc) { labels[i] = bc; ++changed; }
        }
        if (changed == 0) break;
        std::fill(nc.begin(), nc.end(), 0.f);
        std::fill(counts.begin(), counts.end(), 0);
        for (int64_t i = 0; i < n; ++i) {
            float *cp = nc.data() + (int64_t)labels[i] * d;
            ++counts[labels[i]];
            for (int64_t j = 0; j < d; ++j) cp[j] += X[i * d + j];
        }
        for (int64_t c = 0; c < k; ++c) {
            float *cp = nc.data() + c * d;
            if (counts[c] > 0) {
                float inv = 1.f / (float)counts[c];
                for (int64_t j = 0; j < d; ++j) cp[j] *= inv;
            } else {
                std::memcpy(cp, X + (int64_t)pick(rng) * d, d * sizeof(float));
            }
        }
        cent.swap(nc);
    }
    std::memcpy(labels_out, labels.data(), n * sizeof(int32_t));
    std::memcpy(centroids_out, cent.data(), cent.size() * sizeof(float));
}
