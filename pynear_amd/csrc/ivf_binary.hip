// pn_ivfbin: IVFFlatBinaryIndex on the GPU (C-ABI in pynear_amd.h).
//
// Build (BinaryIVF.hpp:127-227), with the reference's random draws made on the host in the same
// order (std::mt19937 + libstdc++ distributions) so the centroids and inverted lists are
// identical to the reference's:
//   * k-means++: per new centroid a kernel folds the Hamming distance to the previous centroid
//     into min_d and emits per-block sums; the host draws `pick`, then `wsel` over the exact
//     integer total, and finds the selected row from the block sums plus one block of min_d.
//   * Lloyd: the assignment is an exact nearest-centroid search (the Hamming scan with k = 1,
//     lower centroid index winning ties like the reference's strict `<`); the majority-vote update
//     counts set bits per cluster over member lists; empty clusters take `pick` rows in cluster order.
//   * inverted lists hold ids in ascending order; their rows are stored contiguously in list order.
// Search (BinaryIVF.hpp:65-111):
//   * coarse: exact top-nprobe centroids per query in (distance, centroid id) order -- the
//     reference's partial_sort over (dist, cid) pairs;
//   * fine, list-major (k <= 32, rows of up to 128 bytes): each inverted list is scanned once for
//     all queries that probe it, then one wave per query merges its per-probe lists (below);
//   * fine, query-major (the other shapes): one workgroup per query scans its probed lists twice. Pass 1 builds a histogram of
//     distances, giving T (the k-th smallest distance) and m = #{d < T}. Pass 2 walks the
//     candidates in scan order (probe rank, then ascending id) with a block-wide prefix count of
//     d <= T, keeping every d < T candidate and the d == T candidates among the first k with
//     d <= T. The result -- all d < T plus the k - m smallest ids of those ties, ascending by
//     (d, id) -- is exactly what the reference's strict-admission max-heap leaves.
#include <algorithm>
#include <cstring>
#include <random>
#include <vector>

#include "common.hpp"
#include "device_util.hpp"
#include "exact_index.hpp"
#include "kernels.hpp"

struct pn_ivfbin {
    int32_t nlist = 256, nprobe = 8, max_iter = 20;
    uint32_t seed = 42;
    int device = 0;
    int64_t n = 0, nbytes = 0, words = 0; // words: padded row width in 32-bit words
    int32_t nc = 0;                        // clusters actually built = min(nlist, n)
    int64_t max_list = 0;                  // rows of the longest inverted list
    pn_exact cent;                         // centroid index (coarse assignment)
    pn::DevBuf lrows, lids, loff, qbuf, probes, pdist, oidx, odist, ocnt;
    pn::DevBuf lm;                         // list-major search: buckets, work items, per-probe lists
    std::vector<int32_t> labels;           // final labels (host), for export
};

namespace pn {

constexpr int kBlockRows = 1024;
constexpr int kMaxIvfK = 2048;

// min_d[i] = min(min_d[i], popcount(x_i ^ c)); block b of kBlockRows rows writes its sum.
__global__ __launch_bounds__(256) void kpp_update_kernel(const uint32_t *__restrict__ X, int64_t n, int words,
                                                         const uint32_t *__restrict__ c, int64_t *__restrict__ min_d,
                                                         int64_t *__restrict__ block_sum) {
    __shared__ int64_t red[256];
    const int64_t base = (int64_t)blockIdx.x * kBlockRows;
    int64_t acc = 0;
    for (int r = threadIdx.x; r < kBlockRows; r += 256) {
        const int64_t i = base + r;
        if (i < n) {
            uint32_t d = 0;
            for (int w = 0; w < words; ++w) d += __popc(X[i * words + w] ^ c[w]);
            int64_t v = min(min_d[i], (int64_t)d);
            min_d[i] = v;
            acc += v;
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_sum[blockIdx.x] = red[0];
}

__global__ void fill_i64_kernel(int64_t *p, int64_t n, int64_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// changed += (new != old); old = new; counts[new] += 1
__global__ void relabel_kernel(const int32_t *__restrict__ nl, int32_t *__restrict__ ol, int64_t n,
                               unsigned long long *__restrict__ changed, int32_t *__restrict__ counts) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool ch = false;
    if (i < n) {
        int32_t v = nl[i];
        ch = v != ol[i];
        ol[i] = v;
        atomicAdd(&counts[v], 1);
    }
    unsigned long long b = __ballot(ch);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(changed, (unsigned long long)__popcll(b));
}

__global__ void scatter_members_kernel(const int32_t *__restrict__ labels, int64_t n, int32_t *__restrict__ cursor,
                                       int32_t *__restrict__ members) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) members[atomicAdd(&cursor[labels[i]], 1)] = (int32_t)i;
}

// Majority vote per bit for cluster blockIdx.x: bit set iff count > members / 2 (integer division).
__global__ __launch_bounds__(256) void majority_kernel(const uint32_t *__restrict__ X, int words,
                                                       const int32_t *__restrict__ members,
                                                       const int32_t *__restrict__ off, uint32_t *__restrict__ C) {
    const int c = blockIdx.x;
    const int b0 = off[c], b1 = off[c + 1];
    const int cnt = b1 - b0;
    if (cnt == 0) return; // empty clusters are reseeded by the host
    const int half = cnt / 2;
    const int nbits = words * 32;
    for (int bit0 = 0; bit0 < nbits; bit0 += 256) {
        const int bit = bit0 + threadIdx.x;
        int ones = 0;
        if (bit < nbits) {
            const int w = bit >> 5, sh = bit & 31;
            for (int j = b0; j < b1; ++j) ones += (X[(int64_t)members[j] * words + w] >> sh) & 1u;
        }
        unsigned long long m = __ballot(bit < nbits && ones > half);
        const int lane = threadIdx.x & 63;
        if (bit < nbits && (lane == 0 || lane == 32)) C[(int64_t)c * words + (bit >> 5)] = (uint32_t)(m >> lane);
    }
}

__global__ void gather_rows_kernel(const uint32_t *__restrict__ X, int words, const int32_t *__restrict__ ids,
                                   int64_t n, uint32_t *__restrict__ out) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n * words) {
        int64_t j = t / words, w = t % words;
        out[t] = X[(int64_t)ids[j] * words + w];
    }
}

// ---------------------------------------------------------------- fine search
template <int W>
__device__ __forceinline__ uint32_t ham_row(const uint32_t *__restrict__ row, const uint32_t *qv, int words) {
    uint32_t d = 0;
    if (W > 0) {
#pragma unroll
        for (int w = 0; w < W; ++w) d = bcnt_acc(row[w] ^ qv[w], d);
    } else {
        for (int w = 0; w < words; ++w) d += __popc(row[w] ^ qv[w]);
    }
    return d;
}

template <int W>
__global__ __launch_bounds__(256) void ivfbin_fine_kernel(const uint32_t *__restrict__ lrows,
                                                          const int32_t *__restrict__ lids,
                                                          const int64_t *__restrict__ loff,
                                                          const int32_t *__restrict__ probes, int np,
                                                          const uint32_t *__restrict__ qs, int words, int nbins,
                                                          int k, int32_t *__restrict__ out_idx,
                                                          int32_t *__restrict__ out_dist, int32_t *__restrict__ out_cnt) {
    extern __shared__ unsigned char smem_raw[];
    uint32_t *hist = (uint32_t *)smem_raw;                                   // nbins
    uint64_t *lt = (uint64_t *)(smem_raw + ((nbins * 4 + 7) & ~7));         // k: (d << 32 | id), d < T
    uint64_t *fin = lt + k;                                                  // k: final keys
    int32_t *tie = (int32_t *)(fin + k);                                     // k: ids with d == T in F
    __shared__ int s_T, s_m, s_M, s_nlt, s_ntie, s_done;
    __shared__ int s_wave[4];

    const int64_t q = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t *qv = qs + q * (int64_t)words; // uniform across the workgroup
    const int32_t *pr = probes + q * (int64_t)np;

    for (int t = tid; t < nbins; t += 256) hist[t] = 0;
    if (tid == 0) {
        s_nlt = 0;
        s_ntie = 0;
        s_done = 0;
    }
    __syncthreads();
    // pass 1: histogram of distances over all probed candidates
    for (int p = 0; p < np; ++p) {
        const int c = pr[p];
        const int64_t b0 = loff[c], b1 = loff[c + 1];
        for (int64_t j = b0 + tid; j < b1; j += 256) atomicAdd(&hist[ham_row<W>(lrows + j * words, qv, words)], 1u);
    }
    __syncthreads();
    if (tid == 0) {
        int M = 0;
        for (int t = 0; t < nbins; ++t) M += (int)hist[t];
        int T = nbins, m = M; // fewer than k+1 candidates: everything is kept
        if (M > k) {
            int cum = 0;
            for (int t = 0; t < nbins; ++t) {
                if (cum + (int)hist[t] >= k) {
                    T = t;
                    m = cum;
                    break;
                }
                cum += (int)hist[t];
            }
        }
        s_T = T;
        s_m = m;
        s_M = M;
    }
    __syncthreads();
    const int T = s_T, m = s_m, M = s_M;
    // pass 2: scan order, prefix rank among d <= T
    int rank_base = 0;
    for (int p = 0; p < np && !s_done; ++p) {
        const int c = pr[p];
        const int64_t b0 = loff[c], b1 = loff[c + 1];
        for (int64_t j0 = b0; j0 < b1; j0 += 256) {
            const int64_t j = j0 + tid;
            const bool valid = j < b1;
            const uint32_t d = valid ? ham_row<W>(lrows + j * words, qv, words) : 0xFFFFFFFFu;
            const bool le = valid && (int)d <= T;
            const unsigned long long bal = __ballot(le);
            const int lane_pre = __popcll(bal & ((1ull << lane) - 1ull));
            if (lane == 0) s_wave[wid] = __popcll(bal);
            __syncthreads();
            int wave_pre = 0, tot = 0;
            for (int w = 0; w < 4; ++w) {
                wave_pre += w < wid ? s_wave[w] : 0;
                tot += s_wave[w];
            }
            const int rank = rank_base + wave_pre + lane_pre;
            if (le) {
                const int32_t id = lids[j];
                if ((int)d < T) {
                    lt[atomicAdd(&s_nlt, 1)] = ((uint64_t)d << 32) | (uint32_t)id;
                } else if (rank < k) {
                    tie[atomicAdd(&s_ntie, 1)] = id;
                }
            }
            rank_base += tot;
            __syncthreads();
            if (tid == 0 && rank_base >= k && s_nlt == m) s_done = 1; // F complete and every d < T seen
            __syncthreads();
            if (s_done) break;
        }
    }
    __syncthreads();
    // keep the (k - m) smallest tie ids (when M > k), merge with the d < T candidates
    const int ntie = s_ntie, nlt = s_nlt;
    const int keep = M > k ? k - m : 0;
    for (int i = tid; i < nlt; i += 256) fin[i] = lt[i];
    for (int i = tid; i < ntie; i += 256) {
        const int32_t id = tie[i];
        int r = 0;
        for (int j = 0; j < ntie; ++j) r += tie[j] < id;
        if (r < keep) fin[nlt + r] = ((uint64_t)(uint32_t)T << 32) | (uint32_t)id;
    }
    __syncthreads();
    const int cnt = nlt + min(keep, ntie);
    // rank sort of the final keys (unique ids) and output, ascending (d, id)
    for (int i = tid; i < cnt; i += 256) {
        const uint64_t key = fin[i];
        int r = 0;
        for (int j = 0; j < cnt; ++j) r += fin[j] < key;
        out_idx[q * k + r] = (int32_t)(uint32_t)(key & 0xFFFFFFFFu);
        out_dist[q * k + r] = (int32_t)(key >> 32);
    }
    for (int i = cnt + tid; i < k; i += 256) {
        out_idx[q * k + i] = -1;
        out_dist[q * k + i] = -1;
    }
    if (tid == 0) out_cnt[q] = cnt;
}

static size_t fine_smem(int nbins, int k) { return ((size_t)nbins * 4 + 7) / 8 * 8 + (size_t)k * 20; }

// ---------------------------------------------------------------- list-major fine search
// The query-major kernel above gives each query one workgroup that walks its probed lists twice.
// The list-major search instead scans each inverted list once for all the queries that probe it:
//   1. bucket: the (query, probe rank) pairs are counted and scattered per list, and every list
//      gets ceil(pairs / 64) work items of 64 pairs;
//   2. scan (one workgroup of four waves per item): lane = one (query, rank) pair with the query's
//      words in VGPRs, each wave streaming a quarter of the list's rows through SGPRs; each lane
//      keeps the K smallest (distance, position) keys -- position = row position in the list
//      storage, ascending id -- and wave 0 merges the four quarters' lists through LDS;
//   3. merge (one wave per query): from the n_probe lists of K keys, the reference heap's result.
// Why the merge is exact (with T the k-th smallest distance, m = #{d < T} < k): every d < T
// candidate sits in its list's K smallest; the reference keeps all of them plus, among the
// candidates F = the first k with d <= T in scan order (probe rank, then position), the k - m
// smallest ids of those with d == T. A list's members of F are its first d <= T rows by position;
// its ties among them that can be kept are its earliest (= lowest-id) ties, at most k - m <= K - m_p
// of them, and the list's K smallest hold its first K - m_p ties. The count of d <= T rows a list
// adds to F is exact from its K keys unless all K are <= T, in which case F ends inside it.
constexpr int kLmGroup = 64; // (query, rank) pairs per work item

__global__ void zero_i32_kernel(int32_t *p, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}

__global__ void lm_count_kernel(const int32_t *__restrict__ probes, int64_t npairs, int32_t *__restrict__ cnt) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < npairs) atomicAdd(&cnt[probes[t]], 1);
}

// Two-level bucketing for nc <= kLmLdsLists: a workgroup counts its kLmChunk pairs per list in
// LDS, then takes one range per list from the global counts (wgb[block][list] = its start within
// the list's bucket); the scatter places each pair at boff + wgb + its rank from an LDS counter.
// (One global atomic per workgroup and list instead of one per pair.)
constexpr int kLmLdsLists = 8192;
constexpr int kLmChunk = 8192;
__global__ __launch_bounds__(256) void lm_count_lds_kernel(const int32_t *__restrict__ probes, int64_t npairs, int nc,
                                                           int32_t *__restrict__ cnt, int32_t *__restrict__ wgb) {
    __shared__ int32_t hist[kLmLdsLists];
    for (int c = threadIdx.x; c < nc; c += 256) hist[c] = 0;
    __syncthreads();
    const int64_t t0 = (int64_t)blockIdx.x * kLmChunk, t1 = min(npairs, t0 + kLmChunk);
    for (int64_t t = t0 + threadIdx.x; t < t1; t += 256) atomicAdd(&hist[probes[t]], 1);
    __syncthreads();
    int32_t *my = wgb + (int64_t)blockIdx.x * nc;
    for (int c = threadIdx.x; c < nc; c += 256) my[c] = hist[c] ? atomicAdd(&cnt[c], hist[c]) : 0;
}

__global__ __launch_bounds__(256) void lm_scatter_lds_kernel(const int32_t *__restrict__ probes, int64_t npairs, int nc,
                                                             const int64_t *__restrict__ boff,
                                                             const int32_t *__restrict__ wgb,
                                                             int32_t *__restrict__ bucket) {
    __shared__ int32_t cur[kLmLdsLists];
    for (int c = threadIdx.x; c < nc; c += 256) cur[c] = 0;
    __syncthreads();
    const int32_t *my = wgb + (int64_t)blockIdx.x * nc;
    const int64_t t0 = (int64_t)blockIdx.x * kLmChunk, t1 = min(npairs, t0 + kLmChunk);
    for (int64_t t = t0 + threadIdx.x; t < t1; t += 256) {
        const int c = probes[t];
        bucket[boff[c] + my[c] + atomicAdd(&cur[c], 1)] = (int32_t)t;
    }
}

// One workgroup of 1024: exclusive scans of the per-list pair counts (boff) and item counts
// (ioff); n_items = total items.
__global__ __launch_bounds__(1024) void lm_scan_kernel(const int32_t *__restrict__ cnt, int nc,
                                                       int64_t *__restrict__ boff, int32_t *__restrict__ ioff,
                                                       int32_t *__restrict__ n_items) {
    __shared__ int64_t sb[1024];
    __shared__ int32_t si[1024];
    __shared__ int64_t carry_b;
    __shared__ int32_t carry_i;
    const int t = threadIdx.x;
    if (t == 0) {
        carry_b = 0;
        carry_i = 0;
    }
    __syncthreads();
    for (int base = 0; base < nc; base += 1024) {
        const int c = base + t;
        const int32_t v = c < nc ? cnt[c] : 0;
        sb[t] = v;
        si[t] = (v + kLmGroup - 1) / kLmGroup;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) { // inclusive Hillis-Steele scan
            const int64_t xb = t >= off ? sb[t - off] : 0;
            const int32_t xi = t >= off ? si[t - off] : 0;
            __syncthreads();
            sb[t] += xb;
            si[t] += xi;
            __syncthreads();
        }
        if (c < nc) {
            boff[c] = carry_b + sb[t] - v;
            ioff[c] = carry_i + si[t] - (v + kLmGroup - 1) / kLmGroup;
        }
        __syncthreads();
        if (t == 1023) {
            carry_b += sb[1023];
            carry_i += si[1023];
        }
        __syncthreads();
    }
    if (t == 0) {
        boff[nc] = carry_b;
        ioff[nc] = carry_i;
        n_items[0] = carry_i;
    }
}

__global__ void lm_scatter_kernel(const int32_t *__restrict__ probes, int64_t npairs,
                                  const int64_t *__restrict__ boff, int32_t *__restrict__ cursor,
                                  int32_t *__restrict__ bucket) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs) return;
    const int c = probes[t];
    bucket[boff[c] + atomicAdd(&cursor[c], 1)] = (int32_t)t; // t = query * n_probe + rank
}

// items[ioff[c] + i] = (c, 64 i) for the lists with pairs
__global__ void lm_items_kernel(const int64_t *__restrict__ boff, const int32_t *__restrict__ ioff, int nc,
                                int64_t *__restrict__ items) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int32_t ni = ioff[c + 1] - ioff[c];
    for (int32_t i = 0; i < ni; ++i) items[ioff[c] + i] = ((int64_t)c << 32) | (int64_t)(i * kLmGroup);
}

template <int W, int K, int RU>
__global__ __launch_bounds__(256) void ivfbin_list_scan_kernel(const uint32_t *__restrict__ lrows,
                                                               const int64_t *__restrict__ loff,
                                                               const int64_t *__restrict__ boff,
                                                               const int32_t *__restrict__ bucket,
                                                               const int64_t *__restrict__ items,
                                                               const int32_t *__restrict__ n_items,
                                                               const uint32_t *__restrict__ qs, int np,
                                                               uint64_t *__restrict__ part) {
    // one work item per workgroup: its 64 (query, rank) pairs on the lanes of every wave, the
    // list's rows split into four contiguous quarters, one per wave; wave 0 merges the four lists
    __shared__ uint32_t red[3][K][64];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t it = blockIdx.x;
    if (it >= n_items[0]) return; // (whole workgroup)
    const int64_t item = items[it];
    const int c = __builtin_amdgcn_readfirstlane((int)(item >> 32));
    const int s = __builtin_amdgcn_readfirstlane((int)(item & 0xFFFFFFFF));
    const int64_t b0 = boff[c], nb = boff[c + 1] - b0;
    const bool live = s + lane < nb;
    const int32_t e = live ? bucket[b0 + s + lane] : 0;
    const int64_t q = e / np;
    uint32_t qv[W];
#pragma unroll
    for (int w = 0; w < W; ++w) qv[w] = live ? qs[q * W + w] : 0u;
    const int64_t r0 = loff[c];
    const int nall = (int)(loff[c + 1] - r0);
    const int quarter = (nall + 3) / 4;
    const int lr0 = min(nall, wv * quarter), lr1 = min(nall, lr0 + quarter);
    uint32_t L[K];
#pragma unroll
    for (int j = 0; j < K; ++j) L[j] = kKey32None;
    uint32_t thr = kKey32None >> kHamIdxBits;
    const uint32_t *sdb = lrows + r0 * W;
    uint32_t R[RU][W];
    for (int lr = lr0; lr < lr1; lr += RU) {
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const int rr = min(lr + u, lr1 - 1); // (clamped rows are loaded, never inserted)
#pragma unroll
            for (int w = 0; w < W; ++w) R[u][w] = sdb[rr * W + w];
        }
        __builtin_amdgcn_sched_barrier(0);
        const int cnt = min(RU, lr1 - lr);
        uint32_t d[RU];
        uint32_t mn = kKey32None;
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            uint32_t acc = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) acc = bcnt_acc(qv[w] ^ R[u][w], acc);
            d[u] = u < cnt ? acc : kKey32None;
            mn = min(mn, d[u]);
        }
        if (mn < thr) {
#pragma unroll
            for (int u = 0; u < RU; ++u) {
                if (d[u] < thr) {
                    list_insert_u32<K>(L, (d[u] << kHamIdxBits) | (uint32_t)(lr + u));
                    thr = L[K - 1] >> kHamIdxBits;
                }
            }
        }
    }
    // keys hold the row's position in the whole list, so the four lists merge by key alone
    if (wv > 0) {
#pragma unroll
        for (int j = 0; j < K; ++j) red[wv - 1][j][lane] = L[j];
    }
    __syncthreads();
    if (wv > 0 || !live) return;
#pragma unroll
    for (int o = 0; o < 3; ++o)
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const uint32_t key = red[o][j][lane];
            if (key < L[K - 1]) list_insert_u32<K>(L, key);
        }
    uint64_t *out = part + (int64_t)e * K;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const uint32_t key = L[j];
        out[j] = key == kKey32None ? kKey64None
                                   : ((uint64_t)(key >> kHamIdxBits) << 32) | (uint64_t)(r0 + (key & kHamIdxMask));
    }
}

// One workgroup of one wave per query: the reference heap's result from the query's n_probe lists
// of K keys (d << 32 | storage position), np * K <= 64 E. Every branch below is uniform across
// the wave, so the barriers only order its lanes' LDS accesses.
template <int E>
__global__ __launch_bounds__(64) void ivfbin_lm_merge_kernel(const uint64_t *__restrict__ part, int64_t nq, int np,
                                                             int K, int k, const int32_t *__restrict__ lids,
                                                             int32_t *__restrict__ out_idx,
                                                             int32_t *__restrict__ out_dist,
                                                             int32_t *__restrict__ out_cnt) {
    __shared__ uint64_t ent[1][64 * E];
    __shared__ uint64_t fin[1][64];
    __shared__ int32_t tie[1][64];
    constexpr int wv = 0;
    const int lane = threadIdx.x;
    const int64_t q = blockIdx.x;
    if (q >= nq) return;
    const int total = np * K;
    const uint64_t *src = part + q * total;
    uint64_t v[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const int i = lane + 64 * s;
        v[s] = i < total ? src[i] : kKey64None;
        ent[wv][i] = v[s];
    }
    __syncthreads();
    auto wsum = [&](int x) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        return x;
    };
    auto count_le = [&](uint32_t t) { // valid entries with d <= t
        int c = 0;
#pragma unroll
        for (int s = 0; s < E; ++s) c += v[s] != kKey64None && (uint32_t)(v[s] >> 32) <= t;
        return wsum(c);
    };
    int nvalid = 0;
#pragma unroll
    for (int s = 0; s < E; ++s) nvalid += v[s] != kKey64None;
    nvalid = wsum(nvalid);
    uint32_t T = 0xFFFFFFFFu; // everything is kept when there are at most k candidates
    int m = nvalid;
    if (nvalid > k) {
        uint32_t lo = 0, hi = 1024; // distances of rows of <= 32 words are <= 1024
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (count_le(mid) >= k) hi = mid;
            else lo = mid + 1;
        }
        T = lo;
        m = T > 0 ? count_le(T - 1) : 0;
    }
    // the d < T candidates (all d when everything is kept) go to fin[0 .. m)
    int nf = 0;
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const bool take = v[s] != kKey64None && (uint32_t)(v[s] >> 32) < T;
        const unsigned long long bal = __ballot(take);
        if (take) {
            const int slot = nf + __popcll(bal & ((1ull << lane) - 1ull));
            const int32_t id = lids[(int64_t)(v[s] & 0xFFFFFFFFull)];
            fin[wv][slot] = (v[s] & 0xFFFFFFFF00000000ull) | (uint32_t)id;
        }
        nf += __popcll(bal);
    }
    int cnt = nf;
    if (nvalid > k) {
        // ties: walk the lists in probe order, F = the first k candidates with d <= T
        int cum = 0, ntie = 0;
        for (int p = 0; p < np && cum < k; ++p) {
            const uint64_t key = lane < K ? ent[wv][p * K + lane] : kKey64None;
            const bool le = key != kKey64None && (uint32_t)(key >> 32) <= T;
            const uint32_t pos = (uint32_t)(key & 0xFFFFFFFFull);
            const int cp = __popcll(__ballot(le));
            const int qp = cp == K ? k - cum : min(cp, k - cum);
            int rank = 0;
            for (int i = 0; i < K; ++i) {
                const bool lei = __shfl(le ? 1 : 0, i, 64) != 0;
                const uint32_t posi = __shfl(pos, i, 64);
                rank += lei && posi < pos;
            }
            const bool tk = le && rank < qp && (uint32_t)(key >> 32) == T;
            const unsigned long long bal = __ballot(tk);
            if (tk) tie[wv][ntie + __popcll(bal & ((1ull << lane) - 1ull))] = lids[(int64_t)pos];
            ntie += __popcll(bal);
            cum += qp;
        }
        __syncthreads();
        // the k - m lowest ids among the ties
        const int keep = k - m;
        if (lane < ntie) {
            const int32_t id = tie[wv][lane];
            int r = 0;
            for (int i = 0; i < ntie; ++i) r += tie[wv][i] < id;
            if (r < keep) fin[wv][m + r] = ((uint64_t)T << 32) | (uint32_t)id;
        }
        cnt = m + min(keep, ntie);
    }
    __syncthreads();
    // ascending (distance, id): rank of each final key
    if (lane < cnt) {
        const uint64_t key = fin[wv][lane];
        int r = 0;
        for (int i = 0; i < cnt; ++i) r += fin[wv][i] < key;
        out_idx[q * k + r] = (int32_t)(uint32_t)(key & 0xFFFFFFFFull);
        out_dist[q * k + r] = (int32_t)(key >> 32);
    }
    for (int i = cnt + lane; i < k; i += 64) {
        out_idx[q * k + i] = -1;
        out_dist[q * k + i] = -1;
    }
    if (lane == 0) out_cnt[q] = cnt;
}

// K >= k keys per (query, probe) list; the insert into a lane's list costs K operations, so K
// follows k closely (10 and 20 for the common k = 10 and 20)
static int lm_list_size(int k) {
    return k <= 4 ? 4 : (k <= 8 ? 8 : (k <= 10 ? 10 : (k <= 16 ? 16 : (k <= 20 ? 20 : 32))));
}

// The list-major path serves k <= 32, rows of 1 .. 32 words, lists shorter than 2^21 rows and
// n_probe * K <= 1024; the query-major kernel takes the rest.
static bool lm_ok(const pn_ivfbin *h, int np, int k) {
    const int w = (int)h->words;
    return tune_get("ivfbin.list_major") != 0 && k <= 32 && (w == 1 || w == 2 || w == 4 || w == 8 || w == 16 || w == 32) &&
           h->max_list < kMaxRowsPerSlice && (int64_t)np * lm_list_size(k) <= 1024;
}

template <int W, int K>
static void launch_lm_scan(const pn_ivfbin *h, const int64_t *boff, const int32_t *bucket, const int64_t *items,
                           const int32_t *n_items, int64_t max_items, const uint32_t *dq, int np, uint64_t *part,
                           hipStream_t st) {
    constexpr int RU = W >= 32 ? 2 : (W >= 16 ? 4 : 8);
    hipLaunchKernelGGL((ivfbin_list_scan_kernel<W, K, RU>), dim3((unsigned)max_items), dim3(256), 0, st,
                       h->lrows.as<uint32_t>(), h->loff.as<int64_t>(), boff, bucket, items, n_items, dq, np, part);
}

template <int W>
static void launch_lm_scan_w(int K, const pn_ivfbin *h, const int64_t *boff, const int32_t *bucket,
                             const int64_t *items, const int32_t *n_items, int64_t max_items, const uint32_t *dq,
                             int np, uint64_t *part, hipStream_t st) {
    switch (K) {
    case 4: launch_lm_scan<W, 4>(h, boff, bucket, items, n_items, max_items, dq, np, part, st); break;
    case 8: launch_lm_scan<W, 8>(h, boff, bucket, items, n_items, max_items, dq, np, part, st); break;
    case 10: launch_lm_scan<W, 10>(h, boff, bucket, items, n_items, max_items, dq, np, part, st); break;
    case 16: launch_lm_scan<W, 16>(h, boff, bucket, items, n_items, max_items, dq, np, part, st); break;
    case 20: launch_lm_scan<W, 20>(h, boff, bucket, items, n_items, max_items, dq, np, part, st); break;
    default: launch_lm_scan<W, 32>(h, boff, bucket, items, n_items, max_items, dq, np, part, st); break;
    }
}

static void launch_fine_lm(pn_ivfbin *h, const uint32_t *dq, int64_t nq, int np, int k, int32_t *oi, int32_t *od,
                           int32_t *oc, hipStream_t st) {
    const int K = lm_list_size(k);
    const int nc = h->nc;
    const int64_t npairs = nq * np;
    const int64_t max_items = npairs / kLmGroup + nc + 1;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t b_cnt = al((size_t)nc * 4), b_boff = al((size_t)(nc + 1) * 8), b_ioff = al((size_t)(nc + 1) * 4),
                 b_ni = al(4), b_bucket = al((size_t)npairs * 4), b_items = al((size_t)max_items * 8),
                 b_part = al((size_t)npairs * K * 8);
    const bool lds = nc <= kLmLdsLists;
    const int64_t nwg = ceil_div(npairs, (int64_t)kLmChunk);
    const size_t b_wgb = lds ? al((size_t)nwg * nc * 4) : 0;
    uint8_t *w = (uint8_t *)h->lm.ensure(2 * b_cnt + b_boff + b_ioff + b_ni + b_bucket + b_items + b_part + b_wgb);
    int32_t *wgb = (int32_t *)(w + 2 * b_cnt + b_boff + b_ioff + b_ni + b_bucket + b_items + b_part);
    int32_t *cnt = (int32_t *)w;
    int32_t *cursor = (int32_t *)(w + b_cnt);
    int64_t *boff = (int64_t *)(w + 2 * b_cnt);
    int32_t *ioff = (int32_t *)(w + 2 * b_cnt + b_boff);
    int32_t *n_items = (int32_t *)(w + 2 * b_cnt + b_boff + b_ioff);
    int32_t *bucket = (int32_t *)(w + 2 * b_cnt + b_boff + b_ioff + b_ni);
    int64_t *items = (int64_t *)(w + 2 * b_cnt + b_boff + b_ioff + b_ni + b_bucket);
    uint64_t *part = (uint64_t *)(w + 2 * b_cnt + b_boff + b_ioff + b_ni + b_bucket + b_items);
    const int32_t *pr = h->probes.as<int32_t>();
    const unsigned gp = (unsigned)ceil_div(npairs, 256);
    {
        StageScope sc(st, STAGE_OTHER);
        hipLaunchKernelGGL(zero_i32_kernel, dim3((unsigned)ceil_div(2 * b_cnt / 4, 256)), dim3(256), 0, st, cnt,
                           (int64_t)(2 * b_cnt / 4));
        if (lds)
            hipLaunchKernelGGL(lm_count_lds_kernel, dim3((unsigned)nwg), dim3(256), 0, st, pr, npairs, nc, cnt, wgb);
        else
            hipLaunchKernelGGL(lm_count_kernel, dim3(gp), dim3(256), 0, st, pr, npairs, cnt);
        hipLaunchKernelGGL(lm_scan_kernel, dim3(1), dim3(1024), 0, st, cnt, nc, boff, ioff, n_items);
        if (lds)
            hipLaunchKernelGGL(lm_scatter_lds_kernel, dim3((unsigned)nwg), dim3(256), 0, st, pr, npairs, nc, boff, wgb,
                               bucket);
        else
            hipLaunchKernelGGL(lm_scatter_kernel, dim3(gp), dim3(256), 0, st, pr, npairs, boff, cursor, bucket);
        hipLaunchKernelGGL(lm_items_kernel, dim3((unsigned)ceil_div(nc, 256)), dim3(256), 0, st, boff, ioff, nc,
                           items);
        PN_HIP(hipGetLastError());
    }
    {
        StageScope sc(st, STAGE_SCAN);
        switch (h->words) {
        case 1: launch_lm_scan_w<1>(K, h, boff, bucket, items, n_items, max_items, dq, np, part, st); break;
        case 2: launch_lm_scan_w<2>(K, h, boff, bucket, items, n_items, max_items, dq, np, part, st); break;
        case 4: launch_lm_scan_w<4>(K, h, boff, bucket, items, n_items, max_items, dq, np, part, st); break;
        case 8: launch_lm_scan_w<8>(K, h, boff, bucket, items, n_items, max_items, dq, np, part, st); break;
        case 16: launch_lm_scan_w<16>(K, h, boff, bucket, items, n_items, max_items, dq, np, part, st); break;
        default: launch_lm_scan_w<32>(K, h, boff, bucket, items, n_items, max_items, dq, np, part, st); break;
        }
        PN_HIP(hipGetLastError());
    }
    StageScope sc_sel(st, STAGE_SELECT);
    const int E = (int)std::max<int64_t>(1, ceil_div((int64_t)np * K, 64));
    const dim3 gm((unsigned)nq);
#define PN_LMM(EV) \
    hipLaunchKernelGGL((ivfbin_lm_merge_kernel<EV>), gm, dim3(64), 0, st, part, nq, np, K, k, h->lids.as<int32_t>(), oi, od, oc)
    if (E <= 1) PN_LMM(1);
    else if (E <= 2) PN_LMM(2);
    else if (E <= 4) PN_LMM(4);
    else if (E <= 8) PN_LMM(8);
    else PN_LMM(16);
#undef PN_LMM
    PN_HIP(hipGetLastError());
}

static void launch_fine(pn_ivfbin *h, const uint32_t *dq, int64_t nq, int np, int k, int32_t *oi, int32_t *od,
                        int32_t *oc, hipStream_t st) {
    if (lm_ok(h, np, k)) {
        launch_fine_lm(h, dq, nq, np, k, oi, od, oc, st);
        return;
    }
    StageScope sc(st, STAGE_SCAN);
    const int nbins = (int)(h->nbytes * 8) + 1;
    const size_t sm = fine_smem(nbins, k);
    PN_REQUIRE(sm <= 64 * 1024, PN_EINVAL, "IVFFlatBinaryIndex search: k or row width too large");
#define PN_FINE(WV)                                                                                       \
    hipLaunchKernelGGL((ivfbin_fine_kernel<WV>), dim3((unsigned)nq), dim3(256), sm, st, h->lrows.as<uint32_t>(), \
                       h->lids.as<int32_t>(), h->loff.as<int64_t>(), h->probes.as<int32_t>(), np, dq,        \
                       (int)h->words, nbins, k, oi, od, oc)
    switch (h->words) {
    case 1: PN_FINE(1); break;
    case 2: PN_FINE(2); break;
    case 4: PN_FINE(4); break;
    case 8: PN_FINE(8); break;
    case 16: PN_FINE(16); break;
    case 32: PN_FINE(32); break;
    default: PN_FINE(0); break;
    }
#undef PN_FINE
    PN_HIP(hipGetLastError());
}

static void ivfbin_build(pn_ivfbin *h, const uint8_t *rows, int64_t n, int64_t nbytes) {
    hipStream_t st = lib_stream();
    bool scan_ok = false;
    const int64_t words = padded_row_elems(PN_METRIC_HAMMING, nbytes, &scan_ok);
    const int32_t k = (int32_t)std::min<int64_t>(h->nlist, n);
    DevBuf X, C, min_d, bsum, nl, ol, counts, cursor, members, offs, chg;
    uint32_t *dX = (uint32_t *)X.ensure((size_t)n * words * 4);
    upload_rows_padded(rows, n, (size_t)nbytes, (size_t)words * 4, dX, st);
    uint32_t *dC = (uint32_t *)C.ensure((size_t)k * words * 4);
    const size_t rowb = (size_t)words * 4;

    // k-means++ (BinaryIVF.hpp:131-159)
    std::mt19937 rng(h->seed);
    std::uniform_int_distribution<size_t> pick(0, (size_t)n - 1);
    PN_HIP(hipMemcpyAsync(dC, dX + (int64_t)pick(rng) * words, rowb, hipMemcpyDeviceToDevice, st));
    const int64_t nblk = ceil_div(n, kBlockRows);
    int64_t *dmin = (int64_t *)min_d.ensure((size_t)n * 8);
    int64_t *dbs = (int64_t *)bsum.ensure((size_t)nblk * 8);
    hipLaunchKernelGGL(fill_i64_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, dmin, n, INT64_MAX);
    std::vector<int64_t> hbs(nblk), hblk(kBlockRows);
    for (int32_t ci = 1; ci < k; ++ci) {
        hipLaunchKernelGGL(kpp_update_kernel, dim3((unsigned)nblk), dim3(256), 0, st, dX, n, (int)words,
                           dC + (int64_t)(ci - 1) * words, dmin, dbs);
        PN_HIP(hipMemcpyAsync(hbs.data(), dbs, nblk * 8, hipMemcpyDeviceToHost, st));
        PN_HIP(hipStreamSynchronize(st));
        int64_t total = 0;
        for (int64_t v : hbs) total += v;
        size_t chosen = pick(rng);
        if (total > 0) {
            std::uniform_int_distribution<int64_t> wsel(0, total - 1);
            const int64_t r = wsel(rng);
            int64_t cum = 0, b = 0;
            while (b < nblk && cum + hbs[b] <= r) cum += hbs[b++]; // first block whose prefix exceeds r
            const int64_t rb0 = b * kBlockRows, rbn = std::min<int64_t>(kBlockRows, n - rb0);
            PN_HIP(hipMemcpyAsync(hblk.data(), dmin + rb0, rbn * 8, hipMemcpyDeviceToHost, st));
            PN_HIP(hipStreamSynchronize(st));
            for (int64_t i = 0; i < rbn; ++i) {
                cum += hblk[i];
                if (cum > r) {
                    chosen = (size_t)(rb0 + i);
                    break;
                }
            }
        }
        PN_HIP(hipMemcpyAsync(dC + (int64_t)ci * words, dX + (int64_t)chosen * words, rowb, hipMemcpyDeviceToDevice,
                              st));
    }

    // Lloyd (BinaryIVF.hpp:161-221)
    int32_t *dnl = (int32_t *)nl.ensure((size_t)n * 4);
    int32_t *dol = (int32_t *)ol.ensure((size_t)n * 4);
    int32_t *dcnt = (int32_t *)counts.ensure((size_t)k * 4);
    int32_t *dcur = (int32_t *)cursor.ensure((size_t)k * 4);
    int32_t *dmem = (int32_t *)members.ensure((size_t)n * 4);
    int32_t *doff = (int32_t *)offs.ensure((size_t)(k + 1) * 4);
    unsigned long long *dchg = (unsigned long long *)chg.ensure(8);
    DevBuf ddist;
    void *dd = ddist.ensure((size_t)n * 4);
    PN_HIP(hipMemsetAsync(dol, 0, (size_t)n * 4, st));
    std::vector<int32_t> hcnt(k), hoff(k + 1);
    pn_exact assign;
    exact_init(&assign, PN_METRIC_HAMMING);
    for (int32_t it = 0; it < h->max_iter; ++it) {
        exact_set_device(&assign, dC, k, nbytes, st);
        exact_search_enqueue(&assign, dX, n, 1, dnl, dd, st);
        PN_HIP(hipMemsetAsync(dchg, 0, 8, st));
        PN_HIP(hipMemsetAsync(dcnt, 0, (size_t)k * 4, st));
        hipLaunchKernelGGL(relabel_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, dnl, dol, n, dchg, dcnt);
        unsigned long long changed = 0;
        PN_HIP(hipMemcpyAsync(&changed, dchg, 8, hipMemcpyDeviceToHost, st));
        PN_HIP(hipMemcpyAsync(hcnt.data(), dcnt, (size_t)k * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipStreamSynchronize(st));
        if (changed == 0) break;
        hoff[0] = 0;
        for (int32_t c = 0; c < k; ++c) hoff[c + 1] = hoff[c] + hcnt[c];
        PN_HIP(hipMemcpyAsync(doff, hoff.data(), (size_t)(k + 1) * 4, hipMemcpyHostToDevice, st));
        PN_HIP(hipMemcpyAsync(dcur, hoff.data(), (size_t)k * 4, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(scatter_members_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, dol, n, dcur,
                           dmem);
        hipLaunchKernelGGL(majority_kernel, dim3((unsigned)k), dim3(256), 0, st, dX, (int)words, dmem, doff, dC);
        for (int32_t c = 0; c < k; ++c) // reseed empty clusters in cluster order
            if (hcnt[c] == 0)
                PN_HIP(hipMemcpyAsync(dC + (int64_t)c * words, dX + (int64_t)pick(rng) * words, rowb,
                                      hipMemcpyDeviceToDevice, st));
        PN_HIP(hipStreamSynchronize(st)); // hoff is reused by the next iteration
    }

    // inverted lists: ids ascending within each list (BinaryIVF.hpp:223-226)
    h->labels.assign(n, 0);
    PN_HIP(hipMemcpyAsync(h->labels.data(), dol, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    PN_HIP(hipStreamSynchronize(st));
    std::vector<int64_t> off(k + 1, 0);
    for (int64_t i = 0; i < n; ++i) ++off[h->labels[i] + 1];
    for (int32_t c = 0; c < k; ++c) off[c + 1] += off[c];
    std::vector<int32_t> ids(n);
    {
        std::vector<int64_t> cur(off.begin(), off.end() - 1);
        for (int64_t i = 0; i < n; ++i) ids[cur[h->labels[i]]++] = (int32_t)i;
    }
    int32_t *dids = (int32_t *)h->lids.ensure((size_t)n * 4);
    int64_t *dloff = (int64_t *)h->loff.ensure((size_t)(k + 1) * 8);
    uint32_t *dl = (uint32_t *)h->lrows.ensure((size_t)n * rowb);
    PN_HIP(hipMemcpyAsync(dids, ids.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
    PN_HIP(hipMemcpyAsync(dloff, off.data(), (size_t)(k + 1) * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)ceil_div(n * words, 256)), dim3(256), 0, st, dX,
                       (int)words, dids, n, dl);
    exact_set_device(&h->cent, dC, k, nbytes, st);
    PN_HIP(hipStreamSynchronize(st));
    h->n = n;
    h->nbytes = nbytes;
    h->words = words;
    h->nc = k;
    h->max_list = 0;
    for (int32_t c = 0; c < k; ++c) h->max_list = std::max<int64_t>(h->max_list, off[c + 1] - off[c]);
}

// Search of nq padded device queries; outputs int32 [nq, k] (-1 padded) and counts.
static void ivfbin_search_enqueue(pn_ivfbin *h, const uint32_t *dq, int64_t nq, int k, int32_t *oi, int32_t *od,
                                  int32_t *oc, hipStream_t st) {
    const int np = std::min(h->nprobe, h->nc);
    if (np <= 0 || h->nc == 0) { // no probes: empty result lists (BinaryIVF.hpp:69-71)
        PN_HIP(hipMemsetAsync(oi, 0xFF, (size_t)nq * k * 4, st));
        PN_HIP(hipMemsetAsync(od, 0xFF, (size_t)nq * k * 4, st));
        PN_HIP(hipMemsetAsync(oc, 0, (size_t)nq * 4, st));
        return;
    }
    int32_t *pr = (int32_t *)h->probes.ensure((size_t)nq * np * 4);
    void *pd = h->pdist.ensure((size_t)nq * np * 4);
    {
        StageScope sc(st, STAGE_COARSE);
        exact_search_enqueue(&h->cent, dq, nq, np, pr, pd, st);
    }
    launch_fine(h, dq, nq, np, k, oi, od, oc, st);
}

} // namespace pn

using namespace pn;

extern "C" pn_status pn_ivfbin_create(int32_t nlist, int32_t nprobe, int32_t max_iter, uint32_t seed,
                                      pn_ivfbin **out) {
    return guard([&] {
        PN_REQUIRE(out, PN_EINVAL, "out is null");
        auto *h = new pn_ivfbin();
        h->nlist = nlist;
        h->nprobe = nprobe;
        h->max_iter = max_iter;
        h->seed = seed;
        h->device = current_device();
        exact_init(&h->cent, PN_METRIC_HAMMING);
        *out = h;
    });
}

extern "C" void pn_ivfbin_destroy(pn_ivfbin *h) {
    if (!h) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(h->device);
    delete h;
    if (prev >= 0) (void)hipSetDevice(prev);
}

extern "C" pn_status pn_ivfbin_set(pn_ivfbin *h, const uint8_t *rows, int64_t n, int64_t nbytes) {
    return guard([&] {
        PN_REQUIRE(h, PN_EINVAL, "index is null");
        PN_REQUIRE(n >= 0 && nbytes >= 0, PN_EINVAL, "negative shape");
        DeviceScope ds(h->device);
        if (n == 0) { // set() on empty data clears the index (BinaryIVF.hpp:50-55)
            h->n = 0;
            h->nbytes = 0;
            h->nc = 0;
            h->labels.clear();
            return;
        }
        PN_REQUIRE(rows && nbytes > 0, PN_EINVAL, "rows must be a non-empty uint8 array");
        PN_REQUIRE(n < (int64_t)INT32_MAX, PN_EINVAL, "too many rows");
        PN_REQUIRE(h->nlist > 0, PN_EINVAL, "nlist must be positive");
        ivfbin_build(h, rows, n, nbytes);
    });
}

extern "C" pn_status pn_ivfbin_info(const pn_ivfbin *h, int64_t *n, int64_t *nbytes, int32_t *n_clusters) {
    return guard([&] {
        PN_REQUIRE(h && n && nbytes && n_clusters, PN_EINVAL, "null argument");
        *n = h->n;
        *nbytes = h->nbytes;
        *n_clusters = h->nc;
    });
}

extern "C" pn_status pn_ivfbin_get_params(const pn_ivfbin *h, int32_t *nlist, int32_t *nprobe) {
    return guard([&] {
        PN_REQUIRE(h && nlist && nprobe, PN_EINVAL, "null argument");
        *nlist = h->nlist;
        *nprobe = h->nprobe;
    });
}

extern "C" pn_status pn_ivfbin_set_nprobe(pn_ivfbin *h, int32_t nprobe) {
    return guard([&] {
        PN_REQUIRE(h, PN_EINVAL, "index is null");
        h->nprobe = nprobe;
    });
}

extern "C" pn_status pn_ivfbin_export(const pn_ivfbin *h, uint8_t *centroids, int32_t *labels) {
    return guard([&] {
        PN_REQUIRE(h, PN_EINVAL, "index is null");
        if (h->n == 0) return;
        DeviceScope ds(h->device);
        if (labels) std::memcpy(labels, h->labels.data(), (size_t)h->n * 4);
        if (centroids) {
            std::vector<uint32_t> c((size_t)h->nc * h->words);
            PN_HIP(hipMemcpy(c.data(), h->cent.db.p, c.size() * 4, hipMemcpyDeviceToHost));
            for (int32_t i = 0; i < h->nc; ++i)
                std::memcpy(centroids + (size_t)i * h->nbytes, c.data() + (size_t)i * h->words, h->nbytes);
        }
    });
}

extern "C" pn_status pn_ivfbin_search(pn_ivfbin *h, const uint8_t *queries, int64_t nq, int64_t nbytes, int32_t k,
                                      int64_t *out_idx, int64_t *out_dist, int32_t *out_count) {
    return guard([&] {
        PN_REQUIRE(h, PN_EINVAL, "index is null");
        PN_REQUIRE(nq >= 0 && k >= 0, PN_EINVAL, "negative size");
        PN_REQUIRE(k <= kMaxIvfK, PN_EINVAL, "IVFFlatBinaryIndex search supports k <= 2048");
        if (nq == 0) return;
        PN_REQUIRE(out_count, PN_EINVAL, "null pointer");
        if (h->n == 0 || k == 0) { // empty index: empty result lists
            std::memset(out_count, 0, (size_t)nq * 4);
            return;
        }
        PN_REQUIRE(nbytes == h->nbytes, PN_EINVAL,
                   "invalid data dimension: index built data and query data dimensions must agree");
        PN_REQUIRE(queries && out_idx && out_dist, PN_EINVAL, "null pointer");
        DeviceScope ds(h->device);
        hipStream_t st = lib_stream();
        const size_t rowb = (size_t)h->words * 4;
        uint32_t *dq = (uint32_t *)h->qbuf.ensure((size_t)nq * rowb);
        upload_rows_padded(queries, nq, (size_t)nbytes, rowb, dq, st);
        int32_t *oi = (int32_t *)h->oidx.ensure((size_t)nq * k * 4);
        int32_t *od = (int32_t *)h->odist.ensure((size_t)nq * k * 4);
        int32_t *oc = (int32_t *)h->ocnt.ensure((size_t)nq * 4);
        ivfbin_search_enqueue(h, dq, nq, k, oi, od, oc, st);
        std::vector<int32_t> hi((size_t)nq * k), hd((size_t)nq * k);
        PN_HIP(hipMemcpyAsync(hi.data(), oi, hi.size() * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipMemcpyAsync(hd.data(), od, hd.size() * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipMemcpyAsync(out_count, oc, (size_t)nq * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipStreamSynchronize(st));
        for (size_t i = 0; i < hi.size(); ++i) {
            out_idx[i] = hi[i];
            out_dist[i] = hd[i];
        }
    });
}

extern "C" pn_status pn_ivfbin_search_device(pn_ivfbin *h, const void *queries, int64_t nq, int32_t k,
                                             int32_t *out_idx, int32_t *out_dist, int32_t *out_count, void *stream) {
    return guard([&] {
        PN_REQUIRE(h, PN_EINVAL, "index is null");
        PN_REQUIRE(nq >= 0 && k > 0 && k <= kMaxIvfK, PN_EINVAL, "bad size");
        PN_REQUIRE(h->n > 0, PN_ERUNTIME, "index is empty");
        if (nq == 0) return;
        DeviceScope ds(h->device);
        hipStream_t st = (hipStream_t)stream;
        const size_t rowb = (size_t)h->words * 4;
        const uint32_t *dq = (const uint32_t *)queries;
        if ((size_t)h->nbytes != rowb) {
            void *tmp = h->qbuf.ensure((size_t)nq * rowb);
            pad_rows_device(queries, nq, (size_t)h->nbytes, rowb, tmp, st);
            dq = (const uint32_t *)tmp;
        }
        ivfbin_search_enqueue(h, dq, nq, k, out_idx, out_dist, out_count, st);
    });
}
