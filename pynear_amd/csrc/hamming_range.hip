// Hamming range search and multi-index-hashing search on the GPU.
//
// Range search replaces BKTree::find / find_batch (BKTree.hpp:71-135, binding
// PythonBindings.cpp:209-230). A BK-tree returns every stored row within the threshold (its walk
// is complete by the triangle inequality); here every row is scored, in the layout of the exact
// Hamming scan: one wave holds 64 queries (one per lane, words in VGPRs) and streams the rows of
// one slice through SGPRs. The output is compressed rows: per query, its matches in ascending row
// order (the reference lists them in breadth-first tree order, which no caller can rely on).
//
//   pass 1  range_scan_kernel<W, RU, false>: lane (query, slice) counts its matches and keeps the
//           first kRangeChunk as (d << 21 | row-in-slice) keys in [slice][chunk][query] order
//           (rows of up to 32 words: d <= 1024 fits the key; wider rows keep none);
//   prefix  range_prefix_kernel: per query, the start of each slice's matches and the total;
//   pass 2  the host turns totals into offsets; range_gather_kernel copies the kept matches to
//           their place, or -- when a (query, slice) had more than kRangeChunk -- the scan runs
//           again and writes every match straight to its place (range_scan_kernel<W, RU, true>).
//
// Multi-index hashing replaces MIHBinaryIndex (MIH.hpp:57-152, binding PythonBindings.cpp:253-272,
// :468-481). The reference looks up, per sub-string table, every key within radius / m bits of the
// query's sub-string and verifies the union of the rows found with a top-k heap. A row is thus a
// candidate exactly when one of its m sub-strings is within radius / m bits of the query's; the
// GPU scores every row with that test and its full distance (the filtered Hamming scan), and keeps
// the k nearest candidates in (distance, row) order.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "device_util.hpp"
#include "exact_index.hpp"
#include "kernels.hpp"

namespace pn {

// W > 0: rows of W words, query words in VGPRs and rows through SGPRs (RU rows per step). W = 0:
// rows of row_words words, query words re-read per row (rows wider than the scan's 32 words).
template <int W, int RU, bool EMIT>
__global__ __launch_bounds__(256) void range_scan_kernel(const uint32_t *__restrict__ db, int64_t n, int64_t rps,
                                                         int64_t row_words, const uint32_t *__restrict__ qs, int64_t nq,
                                                         int64_t nq_pad, uint32_t thr, uint32_t *__restrict__ chunk,
                                                         uint32_t *__restrict__ cnt, const uint32_t *__restrict__ sp,
                                                         const int64_t *__restrict__ qoff, int32_t *__restrict__ oidx,
                                                         int32_t *__restrict__ odist) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t s = blockIdx.y;
    const int64_t r0 = s * rps;
    const int64_t r1 = min(n, r0 + rps);
    const bool live = q < nq;
    constexpr int WV = W > 0 ? W : 1;
    uint32_t qv[WV];
#pragma unroll
    for (int w = 0; w < WV; ++w) qv[w] = (W > 0 && live) ? qs[q * W + w] : 0u;
    int64_t base = 0;
    if (EMIT && live) base = qoff[q] + sp[s * nq_pad + q];
    uint32_t c = 0;
    const int nrows = (int)(r1 - r0);
    auto take = [&](uint32_t d, int lr) {
        if (live && d <= thr) {
            if (EMIT) {
                oidx[base + c] = (int32_t)(r0 + lr);
                odist[base + c] = (int32_t)d;
            } else if (W > 0 && c < (uint32_t)kRangeChunk) { // (wide rows: d may not fit the key)
                chunk[(s * kRangeChunk + c) * nq_pad + q] = (d << kHamIdxBits) | (uint32_t)lr;
            }
            ++c;
        }
    };
    if constexpr (W > 0) {
        const uint32_t *sdb = db + r0 * W;
        uint32_t R[RU][WV];
        for (int lr = 0; lr < nrows; lr += RU) {
#pragma unroll
            for (int u = 0; u < RU; ++u) {
                const int rr = min(lr + u, nrows - 1); // (clamped rows are loaded, never taken)
#pragma unroll
                for (int w = 0; w < W; ++w) R[u][w] = sdb[rr * W + w];
            }
            __builtin_amdgcn_sched_barrier(0);
            const int cntu = min(RU, nrows - lr);
#pragma unroll
            for (int u = 0; u < RU; ++u) {
                uint32_t acc = 0;
#pragma unroll
                for (int w = 0; w < W; ++w) acc = bcnt_acc(qv[w] ^ R[u][w], acc);
                if (u < cntu) take(acc, lr + u);
            }
        }
    } else {
        const uint32_t *qrow = qs + (live ? q : 0) * row_words;
        for (int lr = 0; lr < nrows; ++lr) {
            const uint32_t *row = db + (r0 + lr) * row_words;
            uint32_t acc = 0;
            for (int64_t w = 0; w < row_words; ++w) acc += __popc(qrow[w] ^ row[w]);
            take(acc, lr);
        }
    }
    if (!EMIT && live) cnt[s * nq_pad + q] = c;
}

// Per query: sp[s] = matches in slices before s, tot = all matches, mx = largest slice count.
__global__ __launch_bounds__(256) void range_prefix_kernel(const uint32_t *__restrict__ cnt, int64_t slices,
                                                           int64_t nq, int64_t nq_pad, uint32_t *__restrict__ sp,
                                                           uint32_t *__restrict__ tot, uint32_t *__restrict__ mx) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    uint32_t run = 0, m = 0;
    for (int64_t s = 0; s < slices; ++s) {
        const uint32_t c = cnt[s * nq_pad + q];
        sp[s * nq_pad + q] = run;
        run += c;
        m = max(m, c);
    }
    tot[q] = run;
    mx[q] = m;
}

// One thread per (slice, query): its kept matches (all of them, every count <= kRangeChunk).
__global__ __launch_bounds__(256) void range_gather_kernel(const uint32_t *__restrict__ chunk,
                                                           const uint32_t *__restrict__ cnt,
                                                           const uint32_t *__restrict__ sp,
                                                           const int64_t *__restrict__ qoff, int64_t slices,
                                                           int64_t rps, int64_t nq, int64_t nq_pad,
                                                           int32_t *__restrict__ oidx, int32_t *__restrict__ odist) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t s = t / nq_pad, q = t - s * nq_pad;
    if (s >= slices || q >= nq) return;
    const uint32_t c = min(cnt[s * nq_pad + q], (uint32_t)kRangeChunk);
    const int64_t base = qoff[q] + sp[s * nq_pad + q];
    for (uint32_t j = 0; j < c; ++j) {
        const uint32_t key = chunk[(s * kRangeChunk + j) * nq_pad + q];
        oidx[base + j] = (int32_t)(s * rps + (key & kHamIdxMask));
        odist[base + j] = (int32_t)(key >> kHamIdxBits);
    }
}

template <int W, bool EMIT>
static void launch_range(const void *db, int64_t n, int64_t row_words, const void *queries, int64_t nq,
                         uint32_t thr, const RangePlan &p, uint32_t *chunk, uint32_t *cnt, const uint32_t *sp,
                         const int64_t *qoff, int32_t *oi, int32_t *od, hipStream_t st) {
    constexpr int RU = W >= 32 ? 2 : (W >= 16 ? 4 : 8);
    dim3 grid((unsigned)(p.nq_pad / 256), (unsigned)p.slices);
    hipLaunchKernelGGL((range_scan_kernel<W, RU, EMIT>), grid, dim3(256), 0, st, (const uint32_t *)db, n,
                       p.rows_per_slice, row_words, (const uint32_t *)queries, nq, p.nq_pad, thr, chunk, cnt, sp, qoff,
                       oi, od);
}

template <bool EMIT>
static void dispatch_range(int words, const void *db, int64_t n, int64_t row_words, const void *queries, int64_t nq,
                           uint32_t thr, const RangePlan &p, uint32_t *chunk, uint32_t *cnt, const uint32_t *sp,
                           const int64_t *qoff, int32_t *oi, int32_t *od, hipStream_t st) {
    PN_REQUIRE(p.rows_per_slice <= kMaxRowsPerSlice && p.nq_pad % 256 == 0, PN_EINVAL, "range search: bad plan");
#define PN_RANGE(WV) launch_range<WV, EMIT>(db, n, row_words, queries, nq, thr, p, chunk, cnt, sp, qoff, oi, od, st)
    switch (words) {
    case 0: PN_RANGE(0); break;
    case 1: PN_RANGE(1); break;
    case 2: PN_RANGE(2); break;
    case 4: PN_RANGE(4); break;
    case 8: PN_RANGE(8); break;
    case 16: PN_RANGE(16); break;
    case 32: PN_RANGE(32); break;
    default: fail(PN_EINVAL, "range search: unsupported row width");
    }
#undef PN_RANGE
    PN_HIP(hipGetLastError());
}

void range_count(const void *db, int64_t n, int words, int64_t row_words, const void *queries, int64_t nq,
                 uint32_t threshold, const RangePlan &p, uint32_t *chunk, uint32_t *cnt, uint32_t *sp, uint32_t *tot,
                 uint32_t *mx, hipStream_t st) {
    dispatch_range<false>(words, db, n, row_words, queries, nq, threshold, p, chunk, cnt, nullptr, nullptr, nullptr,
                          nullptr, st);
    hipLaunchKernelGGL(range_prefix_kernel, dim3((unsigned)ceil_div(nq, 256)), dim3(256), 0, st, cnt, p.slices, nq,
                       p.nq_pad, sp, tot, mx);
    PN_HIP(hipGetLastError());
}

void range_write(const void *db, int64_t n, int words, int64_t row_words, const void *queries, int64_t nq,
                 uint32_t threshold, const RangePlan &p, bool rescan, const uint32_t *chunk, const uint32_t *cnt,
                 const uint32_t *sp, const int64_t *qoff, int32_t *out_idx, int32_t *out_dist, hipStream_t st) {
    if (rescan) {
        dispatch_range<true>(words, db, n, row_words, queries, nq, threshold, p, nullptr, nullptr, sp, qoff, out_idx,
                             out_dist, st);
        return;
    }
    const int64_t items = p.slices * p.nq_pad;
    hipLaunchKernelGGL(range_gather_kernel, dim3((unsigned)ceil_div(items, 256)), dim3(256), 0, st, chunk, cnt, sp,
                       qoff, p.slices, p.rows_per_slice, nq, p.nq_pad, out_idx, out_dist);
    PN_HIP(hipGetLastError());
}

static size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

} // namespace pn

using namespace pn;

extern "C" pn_status pn_exact_range(pn_exact *h, const void *queries, int64_t nq, int64_t dim, int64_t threshold,
                                    int64_t *offsets, int32_t *out_idx, int32_t *out_dist, int64_t capacity,
                                    int64_t *total) {
    return guard([&] {
        PN_REQUIRE(h && offsets && total, PN_EINVAL, "null argument");
        PN_REQUIRE(h->metric == PN_METRIC_HAMMING, PN_EINVAL, "range search is for Hamming indexes");
        PN_REQUIRE(nq >= 0 && capacity >= 0, PN_EINVAL, "negative size");
        *total = 0;
        for (int64_t i = 0; i <= nq; ++i) offsets[i] = 0;
        // an empty index finds nothing (BKTree.hpp:79-81); no distance is below a negative threshold
        if (nq == 0 || h->n == 0 || threshold < 0) return;
        PN_REQUIRE(dim == h->dim, PN_EINVAL,
                   "invalid data dimension: index built data and query data dimensions must agree");
        PN_REQUIRE(queries, PN_EINVAL, "queries is null");
        DeviceScope ds(h->device);
        hipStream_t st = lib_stream();
        const int64_t rw = h->row_elems;
        const int words = h->scan_ok ? (int)rw : 0;
        const uint32_t thr = (uint32_t)std::min<int64_t>(threshold, 32 * rw);
        const RangePlan p = range_plan(h, nq);
        const size_t sq = (size_t)p.slices * p.nq_pad * 4;
        const size_t b_q = al256((size_t)nq * rw * 4), b_chunk = al256(sq * kRangeChunk), b_cnt = al256(sq),
                     b_tot = al256((size_t)nq * 4), b_off = al256((size_t)nq * 8);
        uint8_t *w = (uint8_t *)h->rws.ensure(b_q + b_chunk + 2 * b_cnt + 2 * b_tot + b_off);
        void *dq = w;
        uint32_t *chunk = (uint32_t *)(w + b_q);
        uint32_t *cnt = (uint32_t *)(w + b_q + b_chunk);
        uint32_t *sp = (uint32_t *)(w + b_q + b_chunk + b_cnt);
        uint32_t *tot = (uint32_t *)(w + b_q + b_chunk + 2 * b_cnt);
        uint32_t *mx = (uint32_t *)(w + b_q + b_chunk + 2 * b_cnt + b_tot);
        int64_t *qoff = (int64_t *)(w + b_q + b_chunk + 2 * b_cnt + 2 * b_tot);
        upload_rows_padded(queries, nq, (size_t)dim, (size_t)rw * 4, dq, st);
        {
            StageScope sc(st, STAGE_SCAN);
            range_count(h->db.p, h->n, words, rw, dq, nq, thr, p, chunk, cnt, sp, tot, mx, st);
        }
        std::vector<uint32_t> htot(nq), hmx(nq);
        PN_HIP(hipMemcpyAsync(htot.data(), tot, nq * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipMemcpyAsync(hmx.data(), mx, nq * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipStreamSynchronize(st));
        uint32_t mxall = 0;
        for (int64_t i = 0; i < nq; ++i) {
            offsets[i + 1] = offsets[i] + htot[i];
            mxall = std::max(mxall, hmx[i]);
        }
        *total = offsets[nq];
        if (*total == 0 || *total > capacity) return; // (the caller asks again with room for *total)
        PN_REQUIRE(out_idx && out_dist, PN_EINVAL, "output is null");
        PN_REQUIRE(*total < (int64_t)INT32_MAX * 4, PN_EINVAL, "range search: too many matches for one call");
        PN_HIP(hipMemcpyAsync(qoff, offsets, nq * 8, hipMemcpyHostToDevice, st));
        int32_t *oi = (int32_t *)h->rout.ensure((size_t)*total * 8);
        int32_t *od = oi + *total;
        {
            StageScope sc(st, STAGE_SELECT);
            // the kept matches are enough unless a (slice, query) had more, or the rows are wide
            // (pass one then keeps none)
            const bool rescan = words == 0 || mxall > (uint32_t)kRangeChunk;
            range_write(h->db.p, h->n, words, rw, dq, nq, thr, p, rescan, chunk, cnt, sp, qoff, oi, od, st);
        }
        PN_HIP(hipMemcpyAsync(out_idx, oi, (size_t)*total * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipMemcpyAsync(out_dist, od, (size_t)*total * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipStreamSynchronize(st));
    });
}

// ---------------------------------------------------------------- multi-index hashing
struct pn_mih {
    int32_t m = 8;
    int64_t n = 0, nbytes = 0;
    int sub_bytes = 0, group_words = 0;
    int device = 0;
    pn_exact ex; // rows as m groups of group_words words, one sub-string (zero padded) per group
};

namespace pn {

// rows of nbytes bytes -> rows of m groups of 4 g bytes, sub-string t in group t (zero tail)
static std::vector<uint8_t> mih_layout(const uint8_t *rows, int64_t n, int64_t nbytes, int m, int sb, int g) {
    const int64_t ob = (int64_t)m * 4 * g;
    std::vector<uint8_t> out((size_t)n * ob, 0);
    for (int64_t i = 0; i < n; ++i)
        for (int t = 0; t < m; ++t) std::memcpy(&out[i * ob + t * 4 * g], rows + i * nbytes + t * sb, sb);
    return out;
}

} // namespace pn

extern "C" pn_status pn_mih_create(int32_t m, pn_mih **out) {
    return guard([&] {
        PN_REQUIRE(out, PN_EINVAL, "out is null");
        PN_REQUIRE(m >= 1, PN_EINVAL, "MIH: m must be at least 1");
        auto *h = new pn_mih();
        h->m = m;
        h->device = current_device();
        exact_init(&h->ex, PN_METRIC_HAMMING);
        *out = h;
    });
}

extern "C" void pn_mih_destroy(pn_mih *h) {
    if (!h) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(h->device);
    delete h;
    if (prev >= 0) (void)hipSetDevice(prev);
}

extern "C" pn_status pn_mih_set(pn_mih *h, const uint8_t *rows, int64_t n, int64_t nbytes) {
    return guard([&] {
        PN_REQUIRE(h, PN_EINVAL, "index is null");
        PN_REQUIRE(n >= 0 && nbytes >= 0, PN_EINVAL, "negative shape");
        DeviceScope ds(h->device);
        if (n == 0) { // MIH.hpp:60-63: an empty set clears the index
            h->n = 0;
            PN_REQUIRE(pn_exact_set(&h->ex, nullptr, 0, 0) == PN_OK, PN_ERUNTIME, pn_last_error());
            return;
        }
        PN_REQUIRE(rows, PN_EINVAL, "rows is null");
        PN_REQUIRE(nbytes >= h->m && nbytes % h->m == 0, PN_EINVAL, "MIH: descriptor byte width must be divisible by m");
        const int sb = (int)(nbytes / h->m);
        PN_REQUIRE(sb <= 8, PN_EINVAL,
                   "MIH: sub-string width exceeds 8 bytes (uint64_t key capacity). Increase m so that nbytes/m <= 8.");
        const int g = sb <= 4 ? 1 : 2;
        PN_REQUIRE((int64_t)h->m * g <= 32, PN_EINVAL,
                   "MIH on the GPU takes at most 32 sub-strings of up to 4 bytes or 16 of up to 8 bytes");
        std::vector<uint8_t> lay = mih_layout(rows, n, nbytes, h->m, sb, g);
        h->n = 0;
        PN_REQUIRE(pn_exact_set(&h->ex, lay.data(), n, (int64_t)h->m * 4 * g) == PN_OK, PN_ERUNTIME, pn_last_error());
        h->n = n;
        h->nbytes = nbytes;
        h->sub_bytes = sb;
        h->group_words = g;
    });
}

extern "C" pn_status pn_mih_info(const pn_mih *h, int32_t *m, int64_t *n, int64_t *nbytes) {
    return guard([&] {
        PN_REQUIRE(h && m && n && nbytes, PN_EINVAL, "null argument");
        *m = h->m;
        *n = h->n;
        *nbytes = h->nbytes;
    });
}

extern "C" pn_status pn_mih_search(pn_mih *h, const uint8_t *queries, int64_t nq, int64_t nbytes, int32_t k,
                                   int32_t radius, int64_t *out_idx, int64_t *out_dist, int64_t *out_count) {
    return guard([&] {
        PN_REQUIRE(h, PN_EINVAL, "index is null");
        PN_REQUIRE(h->n > 0, PN_ERUNTIME, "MIH: the index is empty; call set() with a non-empty array first");
        PN_REQUIRE(nq >= 0 && k >= 0, PN_EINVAL, "negative size");
        PN_REQUIRE(radius >= 0, PN_EINVAL, "MIH: radius must be non-negative");
        PN_REQUIRE(nbytes == h->nbytes, PN_EINVAL, "MIH: query byte width must equal the indexed width");
        const int kk = (int)std::min<int64_t>(k, h->n);
        if (nq == 0) return;
        PN_REQUIRE(out_count, PN_EINVAL, "null pointer");
        if (kk == 0) {
            for (int64_t i = 0; i < nq; ++i) out_count[i] = 0;
            return;
        }
        PN_REQUIRE(queries && out_idx && out_dist, PN_EINVAL, "null pointer");
        DeviceScope ds(h->device);
        hipStream_t st = lib_stream();
        pn_exact *ex = &h->ex;
        std::vector<uint8_t> lay = mih_layout(queries, nq, nbytes, h->m, h->sub_bytes, h->group_words);
        void *dq = ex->qbuf.ensure((size_t)nq * ex->row_elems * 4);
        upload_rows_padded(lay.data(), nq, (size_t)h->m * 4 * h->group_words, (size_t)ex->row_elems * 4, dq, st);
        int32_t *oi = (int32_t *)ex->oidx.ensure((size_t)nq * kk * 4);
        int32_t *od = (int32_t *)ex->odist.ensure((size_t)nq * kk * 4);
        SubFilter f;
        f.group_words = h->group_words;
        f.subs = h->m;
        f.rsub = radius / h->m; // MIH.hpp:102
        exact_search_enqueue(ex, dq, nq, kk, oi, od, st, f);
        std::vector<int32_t> hi((size_t)nq * kk), hd((size_t)nq * kk);
        PN_HIP(hipMemcpyAsync(hi.data(), oi, hi.size() * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipMemcpyAsync(hd.data(), od, hd.size() * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipStreamSynchronize(st));
        for (int64_t i = 0; i < nq; ++i) {
            int64_t c = 0;
            for (int j = 0; j < kk; ++j) {
                const size_t e = (size_t)i * kk + j;
                out_idx[e] = hi[e];
                out_dist[e] = hi[e] >= 0 ? hd[e] : -1;
                c += hi[e] >= 0;
            }
            out_count[i] = c;
        }
    });
}
