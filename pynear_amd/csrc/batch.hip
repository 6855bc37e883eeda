// pn_batch: exact L2 searches over many independent indices of uneven sizes, served together.
//
// Every sub-search is pn_exact_search_device on its own index (VPTree.hpp:143-166 semantics:
// min(n, k) neighbours, ascending (distance, index), distances bit-identical to the reference's
// dist_l2_f_avx2). A plan lays the sub-searches out as flat work lists and runs them as
//   1. one split launch (bf16 hi/lo + norms of every query of every index),
//   2. persistent scan launches: all (index, query block, row slice) tiles of all indices in
//      three queues (rows of 128, 192 or 256 padded columns, whose scan holds more registers and
//      stages tiles in LDS; the same for indices of at most 64 queries, whose waves take
//      different slices instead of padding queries; the rest), longest tiles first; each tile
//      runs the same device body as the single search,
//   3. merge launches (one per private-list size), one rerank (with the completeness proof) and one
//      fallback launch over all queries,
// captured once in a HIP graph that every run replays. Slice length, persistent grid size and
// tile order are the plan's schedule; pn_batch_tune measures a set of schedules on the device and
// keeps the fastest (optionally persisted in a text table keyed by the batch's shape).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <numeric>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "device_util.hpp"
#include "exact_index.hpp"
#include "l2_mfma.hpp"
#include "l2_mfma_dev.hpp"
#include "merge_dev.hpp"

namespace pn {

// One sub-search of a batch, as the device kernels see it.
struct BatchEntry {
    const __bf16 *Xh, *Xl;
    const float *xn, *X;
    const unsigned int *xn_max;
    int64_t xstride, n, d;
    int dp, kk, KL, K2, K2c;
    int G; // 32-query groups per scan workgroup: 4, or 2 / 1 for at most 64 / 32 queries
    uint32_t lowm; // row bits of the scan's keys-only lists (rows per slice <= lowm + 1)
    float eps_coef;
    const float *Q;
    int64_t qstride, nq, nq_pad, gx, S, rps;
    __bf16 *Qh, *Ql;
    float *qn;
    uint64_t *part;
    uint32_t *gthr; // per padded query: the bound its slices share (null: not shared)
    int32_t *cidx, *flags;
    float *capx;
    int32_t *out_idx;
    float *out_dist;
    int64_t row_base, q_base; // prefixes over the entries: padded query rows, queries
};

// The entry with every pointer marked as global memory (as_global): kernels take their entry
// through this copy.
__device__ __forceinline__ BatchEntry global_entry(const BatchEntry &src) {
    BatchEntry e = src;
    e.Xh = as_global(e.Xh);
    e.Xl = as_global(e.Xl);
    e.xn = as_global(e.xn);
    e.X = as_global(e.X);
    e.xn_max = as_global(e.xn_max);
    e.Q = as_global(e.Q);
    e.Qh = as_global(e.Qh);
    e.Ql = as_global(e.Ql);
    e.qn = as_global(e.qn);
    e.part = as_global(e.part);
    e.gthr = e.gthr ? as_global(e.gthr) : nullptr;
    e.cidx = as_global(e.cidx);
    e.flags = as_global(e.flags);
    e.capx = as_global(e.capx);
    e.out_idx = as_global(e.out_idx);
    e.out_dist = as_global(e.out_dist);
    return e;
}

struct BatchCfg {
    int slice_rows = 512; // target rows per slice (rounded to 32-row tiles)
    int wg_per_cu = 2;    // persistent scan workgroups per CU
    int order = 1;        // 1: tiles by decreasing cost; 0: index order; 2: by slice, then by cost
};

__device__ __forceinline__ int find_entry(const int64_t *base, int n, int64_t x) {
    int lo = 0, hi = n - 1; // last entry with base[e] <= x
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void batch_split_kernel(const BatchEntry *__restrict__ E, const int64_t *__restrict__ row_base,
                                                          int ne, int64_t rows, int32_t *__restrict__ counter) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (blockIdx.x == 0 && threadIdx.x < 3) counter[threadIdx.x] = 0; // the scan queues of this run
    if (r >= rows) return;
    const BatchEntry e = global_entry(E[find_entry(row_base, ne, r)]);
    split_row(e.Q, e.nq, e.d, e.qstride, e.dp, e.Qh, e.Ql, e.qn, r - e.row_base);
    if (e.gthr && (threadIdx.x & 63) == 0) e.gthr[r - e.row_base] = kKey32None;
}

template <int KL>
__device__ __forceinline__ void batch_tile(uint64_t *lds_raw, const BatchEntry &e, int64_t qb, int64_t s, bool pipe) {
    auto lds = reinterpret_cast<uint64_t (*)[32][KL]>(lds_raw);
    if (pipe) {
        switch (e.dp) {
        case 16: scan_small_pipe<KL, 3, 1>(lds, e.Xh, e.Xl, e.xn, e.n, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.G, e.gthr); return;
        case 32: scan_small_pipe<KL, 3, 2>(lds, e.Xh, e.Xl, e.xn, e.n, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.G, e.gthr); return;
        case 48: scan_small_pipe<KL, 3, 3>(lds, e.Xh, e.Xl, e.xn, e.n, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.G, e.gthr); return;
        case 64: scan_small_pipe<KL, 3, 4>(lds, e.Xh, e.Xl, e.xn, e.n, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.G, e.gthr); return;
        default: break;
        }
    }
    switch (e.dp) {
    case 16: scan_small<KL, 3, 1, true>(lds, e.Xh, e.Xl, e.xn, e.n, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.G, e.gthr); break;
    case 32: scan_small<KL, 3, 2, true>(lds, e.Xh, e.Xl, e.xn, e.n, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.G, e.gthr); break;
    case 48: scan_small<KL, 3, 3, true>(lds, e.Xh, e.Xl, e.xn, e.n, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.G, e.gthr); break;
    case 64: scan_small<KL, 3, 4, true>(lds, e.Xh, e.Xl, e.xn, e.n, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.G, e.gthr); break;
    default:
        scan_generic<KL, 3, true>(lds, e.Xh, e.Xl, e.xn, e.n, e.dp, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.gthr);
    }
}

// Tiles of rows padded to 128, 192 or 256 columns: the scan with register-resident queries.
template <int KL>
__device__ __forceinline__ void batch_tile_wide(uint64_t *lds_raw, __bf16 *stage, const BatchEntry &e, int64_t qb,
                                                int64_t s, bool pipe) {
    auto lds = reinterpret_cast<uint64_t (*)[32][KL]>(lds_raw);
    if (pipe) {
        if (e.dp == 128)
            scan_wide_pipe<KL, 3, 2>(lds, stage, e.Xh, e.Xl, e.xn, e.n, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.gthr);
        else if (e.dp == 192)
            scan_wide_pipe<KL, 3, 3>(lds, stage, e.Xh, e.Xl, e.xn, e.n, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.gthr);
        else
            scan_wide_pipe<KL, 3, 4>(lds, stage, e.Xh, e.Xl, e.xn, e.n, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.gthr);
        return;
    }
    if (e.dp == 128)
        scan_wide<KL, 3, 2, true>(lds, stage, e.Xh, e.Xl, e.xn, e.n, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.gthr);
    else if (e.dp == 192)
        scan_wide<KL, 3, 3, true>(lds, stage, e.Xh, e.Xl, e.xn, e.n, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.gthr);
    else
        scan_wide<KL, 3, 4, true>(lds, stage, e.Xh, e.Xl, e.xn, e.n, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.gthr);
}

// Wide-row tiles of indices with at most 64 queries: the waves scan different slices.
template <int KL>
__device__ __forceinline__ void batch_tile_wide_direct(uint64_t *lds_raw, const BatchEntry &e, int64_t qb, int64_t s) {
    auto lds = reinterpret_cast<uint64_t (*)[32][KL]>(lds_raw);
    if (e.dp == 128)
        scan_wide_direct<KL, 3, 2, true>(lds, e.Xh, e.Xl, e.xn, e.n, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.G, e.gthr);
    else if (e.dp == 192)
        scan_wide_direct<KL, 3, 3, true>(lds, e.Xh, e.Xl, e.xn, e.n, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.G, e.gthr);
    else
        scan_wide_direct<KL, 3, 4, true>(lds, e.Xh, e.Xl, e.xn, e.n, e.Qh, e.Ql, e.qn, e.rps, e.part, e.S * e.KL, qb, s, e.G, e.gthr);
}

static bool wide_dp(int dp) { return dp == 128 || dp == 192 || dp == 256; }

// Persistent: workgroups take tiles (entry, query block, slice group) from one queue until it is
// empty. One launch per tile class, each at its own occupancy: CLS 0 rows of at most 64 padded
// columns (or other widths), CLS 1 rows of 128, 192 or 256 columns (LDS-staged tiles), CLS 2
// those of indices with at most 64 queries (waves on different slices, direct loads).
template <int CLS>
__global__ __launch_bounds__(256) void batch_scan_kernel(const BatchEntry *__restrict__ E,
                                                         const int2 *__restrict__ tiles, int64_t ntiles,
                                                         int32_t *__restrict__ counter, int pipe) {
    __shared__ uint64_t lds_raw[4 * 32 * 32];
    __shared__ __align__(16) __bf16 stage[CLS == 1 ? wide_stage_elems<4, 2>() : 8];
    __shared__ int64_t s_slot;
    for (;;) {
        if (threadIdx.x == 0) s_slot = atomicAdd(counter, 1);
        __syncthreads();
        const int64_t slot = s_slot;
        if (slot >= ntiles) break;
        const int2 t = tiles[slot];
        const BatchEntry e = global_entry(E[t.x]);
        const int64_t qb = t.y / e.S, s = t.y % e.S;
        if (CLS == 1) {
            if (e.KL == 8) batch_tile_wide<8>(lds_raw, stage, e, qb, s, pipe & 1);
            else if (e.KL == 16) batch_tile_wide<16>(lds_raw, stage, e, qb, s, pipe & 1);
            else batch_tile_wide<32>(lds_raw, stage, e, qb, s, pipe & 1);
        } else if (CLS == 2) {
            if (e.KL == 8) batch_tile_wide_direct<8>(lds_raw, e, qb, s);
            else if (e.KL == 16) batch_tile_wide_direct<16>(lds_raw, e, qb, s);
            else batch_tile_wide_direct<32>(lds_raw, e, qb, s);
        } else {
            if (e.KL == 8) batch_tile<8>(lds_raw, e, qb, s, pipe & 2);
            else if (e.KL == 16) batch_tile<16>(lds_raw, e, qb, s, pipe & 2);
            else batch_tile<32>(lds_raw, e, qb, s, pipe & 2);
        }
        __syncthreads(); // LDS and s_slot are reused by the next tile
    }
}

// One launch per private-list size K (the kernel's registers grow with K, and with them its
// occupancy falls): gids lists the batch's global query numbers whose merge takes lists of K.
template <int K>
__global__ __launch_bounds__(256) void batch_merge_kernel(const BatchEntry *__restrict__ E,
                                                          const int64_t *__restrict__ q_base, int ne,
                                                          const int32_t *__restrict__ gids, int64_t count) {
    __shared__ uint64_t wbuf_all[4][256];
    uint64_t *wbuf = wbuf_all[threadIdx.x >> 6];
    const int64_t i = xcd_contiguous_block() * 4 + (threadIdx.x >> 6);
    if (i >= count) return;
    const int64_t g = gids[i];
    const BatchEntry e = global_entry(E[find_entry(q_base, ne, g)]);
    uint64_t unused;
    merge_wave_query<K>(e.part, e.S, e.KL, e.S * e.KL, 1, g - e.q_base, e.K2c, OUT_F32, e.cidx, e.capx, wbuf, unused);
}

// The merge without private lists: one launch for all queries of the batch.
__global__ __launch_bounds__(256) void batch_merge_lean_kernel(const BatchEntry *__restrict__ E,
                                                               const int64_t *__restrict__ q_base, int ne,
                                                               int64_t nq, bool rounds_only) {
    __shared__ uint64_t wbuf_all[4][256];
    const int64_t g = xcd_contiguous_block() * 4 + (threadIdx.x >> 6);
    if (g >= nq) return;
    const BatchEntry e = global_entry(E[find_entry(q_base, ne, g)]);
    uint64_t unused;
    merge_wave_lean(e.part, e.S, e.KL, e.S * e.KL, 1, g - e.q_base, e.K2c, OUT_F32, e.cidx, e.capx,
                    wbuf_all[threadIdx.x >> 6], rounds_only, unused);
}

// The lean merge and the rerank of every query in one launch (the candidates stay in registers).
__global__ __launch_bounds__(256) void batch_merge_rerank_kernel(const BatchEntry *__restrict__ E,
                                                                 const int64_t *__restrict__ q_base, int ne,
                                                                 int64_t nq, bool rounds_only) {
    __shared__ uint64_t wbuf_all[4][256];
    const int64_t g = xcd_contiguous_block() * 4 + (threadIdx.x >> 6);
    if (g >= nq) return;
    const BatchEntry e = global_entry(E[find_entry(q_base, ne, g)]);
    const int64_t q = g - e.q_base;
    uint64_t mine;
    merge_wave_lean<true>(e.part, e.S, e.KL, e.S * e.KL, 1, q, e.K2c, OUT_F32, nullptr, nullptr,
                          wbuf_all[threadIdx.x >> 6], rounds_only, mine);
    rerank_regs<false>(e.X, e.xstride, e.d, e.Q, e.qstride, e.qn, e.xn_max, mine, e.K2c, e.part, e.S, e.KL,
                       e.S * e.KL, e.kk, e.n, e.eps_coef, e.out_idx, e.out_dist, e.flags, q, e.lowm);
}

// a lane that merges one slice holds at most KL keys; with more slices per lane it needs K2c
static int merge_cap(const BatchEntry &e) { return e.S <= 64 ? e.KL : e.K2; }

__global__ __launch_bounds__(256) void batch_rerank_kernel(const BatchEntry *__restrict__ E,
                                                           const int64_t *__restrict__ q_base, int ne, int64_t nq) {
    const int64_t g = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (g >= nq) return;
    const BatchEntry e = global_entry(E[find_entry(q_base, ne, g)]);
    rerank_query<false>(e.X, e.xstride, e.d, e.Q, e.qstride, e.qn, e.xn_max, e.cidx, e.capx, e.K2c, e.part, e.S, e.KL,
                 e.S * e.KL, e.kk, e.n, e.eps_coef, e.out_idx, e.out_dist, e.flags, g - e.q_base, e.lowm);
}

// One workgroup per chunk of kFallbackChunk queries of the batch: their flags are read at once
// (each thread finds its query's entry), and the flagged queries run the exact scan one after
// another (chunks in parallel).
__global__ __launch_bounds__(256) void batch_fallback_kernel(const BatchEntry *__restrict__ E,
                                                             const int64_t *__restrict__ q_base, int ne, int64_t nq) {
    __shared__ uint64_t wres[4][32];
    __shared__ uint32_t s_flagged;
    for (int64_t g0 = (int64_t)blockIdx.x * kFallbackChunk; g0 < nq; g0 += (int64_t)gridDim.x * kFallbackChunk) {
        if (threadIdx.x < 64) { // (wave 0: bit t for query g0 + t)
            bool f = false;
            if (threadIdx.x < kFallbackChunk && g0 + threadIdx.x < nq) {
                const int64_t g = g0 + threadIdx.x;
                const int ei = find_entry(q_base, ne, g);
                f = as_global(E[ei].flags)[g - q_base[ei]] != 0;
            }
            const uint64_t bal = __ballot(f);
            if (threadIdx.x == 0) s_flagged = (uint32_t)bal;
        }
        __syncthreads();
        const uint32_t fl = s_flagged;
        __syncthreads(); // (s_flagged is rewritten for the next chunk)
        for (uint32_t m = fl; m; m &= m - 1) {
            const int64_t g = g0 + __builtin_ctz(m);
            const BatchEntry e = global_entry(E[find_entry(q_base, ne, g)]);
            fallback_query<32>(wres, e.X, e.xstride, e.n, e.d, e.Q, e.qstride, e.kk, e.out_idx, e.out_dist,
                               g - e.q_base);
        }
    }
}

static int kl_of(int kk) { return kk <= 6 ? 8 : (kk <= 12 ? 16 : 32); } // as the single search

} // namespace pn

struct pn_batch {
    int device = 0, cus = 256;
    std::vector<pn_exact *> idx;
    std::vector<uint64_t> gen; // each index's set() generation when the batch was planned
    std::vector<const float *> q;
    std::vector<int64_t> nq;
    std::vector<int32_t> k;
    std::vector<int32_t *> oi;
    std::vector<float *> od;
    pn::BatchCfg cfg;
    int64_t ntiles = 0, ncls[3] = {0, 0, 0}, rows = 0, queries = 0;
    pn::DevBuf entries, row_base, q_base, tiles, counter, scratch, mgids;
    int64_t mcount[4] = {0, 0, 0, 0}; // merge queries per list size 8, 16, 32, 64
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~pn_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
    }
};

namespace pn {

static std::string batch_signature(const pn_batch *b) {
    std::ostringstream s;
    s << "b" << b->idx.size();
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < b->idx.size(); ++i)
        for (int64_t v : {b->idx[i]->n, b->idx[i]->dim, b->nq[i], (int64_t)b->k[i]}) {
            h ^= (uint64_t)v;
            h *= 1099511628211ull;
        }
    s << "_" << std::hex << h;
    return s.str();
}

// Lays out entries, scratch and tiles for cfg, and captures the whole search in a graph.
static void batch_plan(pn_batch *b, const BatchCfg &cfg) {
    const int ne = (int)b->idx.size();
    std::vector<BatchEntry> E(ne);
    std::vector<int64_t> rbase(ne), qbase(ne), off(ne);
    size_t bytes = 0;
    auto take = [&](size_t n) { // 256-byte aligned sub-allocations of one scratch buffer
        const size_t o = bytes;
        bytes += (n + 255) / 256 * 256;
        return o;
    };
    struct Off { size_t qh, ql, qn, part, gthr, cidx, capx, flags; };
    std::vector<Off> O(ne);
    int64_t rows = 0, queries = 0;
    struct TileRec {
        double cost;
        int64_t round; // the tile's slice (or group of slices) within its query block
        int2 t;
    };
    std::vector<TileRec> tc[3]; // tiles of each scan class
    const bool wide_on = tune_get("l2.wide") != 0;
    const bool share = tune_get("batch.share") != 0;
    for (int i = 0; i < ne; ++i) {
        pn_exact *h = b->idx[i];
        BatchEntry &e = E[i];
        e.n = h->n;
        e.d = h->dim;
        e.dp = h->l2.dp;
        e.xstride = h->row_elems;
        e.kk = (int)std::min<int64_t>(b->k[i], h->n);
        e.KL = kl_of(e.kk);
        e.K2 = 2 * e.KL;
        e.K2c = (int)std::min<int64_t>(e.K2, e.n);
        e.eps_coef = 6.103515625e-05f + (float)e.dp * 4.76837158203125e-07f; // three products per term
        e.nq = b->nq[i];
        e.qstride = e.d;
        e.gx = ceil_div(e.nq, (int64_t)128);
        e.nq_pad = e.gx * 128;
        // slices of at most 4096 rows: a keys-only list key keeps at most 12 row bits, so at least
        // 19 bits of the approximation (a relative step of at most 2^-11; a coarser key only
        // widens the rerank's candidate filter and its completeness bound)
        int64_t S = std::max<int64_t>(1, ceil_div(e.n, (int64_t)std::min(4096, std::max(32, cfg.slice_rows))));
        // few queries: fewer 32-query groups per workgroup and its waves on consecutive slices, so
        // no wave multiplies padding queries (S rounded up to whole groups of 4 / G slices; the
        // trailing ones may hold no rows)
        e.G = e.nq <= 32 ? 1 : (e.nq <= 64 ? 2 : 4);
        const int64_t m = 4 / e.G;
        S = ceil_div(S, m) * m;
        e.rps = ceil_div(ceil_div(e.n, S), (int64_t)32) * 32;
        e.S = e.G == 4 ? ceil_div(e.n, e.rps) : S;
        int rb = 5;
        while ((int64_t)1 << rb < e.rps) ++rb;
        e.lowm = (1u << rb) - 1u;
        e.row_base = rows;
        e.q_base = queries;
        rbase[i] = rows;
        qbase[i] = queries;
        rows += e.nq_pad;
        queries += e.nq;
        O[i].qh = take((size_t)e.nq_pad * e.dp * 2);
        O[i].ql = take((size_t)e.nq_pad * e.dp * 2);
        O[i].qn = take((size_t)e.nq_pad * 4);
        O[i].part = take((size_t)e.S * e.KL * e.nq_pad * 8);
        O[i].gthr = take((size_t)e.nq_pad * 4);
        O[i].cidx = take((size_t)e.nq * e.K2c * 4);
        O[i].capx = take((size_t)e.nq * e.K2c * 4);
        O[i].flags = take((size_t)e.nq * 4);
        const double cost = (double)e.rps * (e.dp / 16 + 4); // a tile: rows x (products + keys)
        const int cls = wide_on && wide_dp(e.dp) ? (e.G == 4 ? 1 : 2) : 0;
        // tiles: (query block, slice) for G = 4; (the one query block, group of 4 / G slices) else
        const int64_t ntile = e.G == 4 ? e.gx * e.S : e.S / m;
        for (int64_t t = 0; t < ntile; ++t) tc[cls].push_back({cost, e.G == 4 ? t % e.S : t, make_int2(i, (int)t)});
    }
    // order 1: longest tiles first. Order 2: by slice, then longest first -- a query block's
    // slices run in different rounds of the persistent grid, and the later ones start from the
    // bound the earlier ones published.
    std::vector<TileRec> tl;
    for (int c = 0; c < 3; ++c) {
        if (cfg.order == 1)
            std::stable_sort(tc[c].begin(), tc[c].end(), [](const auto &a, const auto &x) { return a.cost > x.cost; });
        else if (cfg.order == 2)
            std::stable_sort(tc[c].begin(), tc[c].end(), [](const auto &a, const auto &x) {
                return a.round != x.round ? a.round < x.round : a.cost > x.cost;
            });
        tl.insert(tl.end(), tc[c].begin(), tc[c].end());
    }
    char *base = (char *)b->scratch.ensure(bytes);
    for (int i = 0; i < ne; ++i) {
        pn_exact *h = b->idx[i];
        BatchEntry &e = E[i];
        e.Xh = h->l2.hi.as<__bf16>();
        e.Xl = h->l2.lo.as<__bf16>();
        e.xn = h->l2.norm.as<float>();
        e.xn_max = h->l2.norm_max.as<unsigned int>();
        e.X = (const float *)h->db.p;
        e.Q = b->q[i];
        e.Qh = (__bf16 *)(base + O[i].qh);
        e.Ql = (__bf16 *)(base + O[i].ql);
        e.qn = (float *)(base + O[i].qn);
        e.part = (uint64_t *)(base + O[i].part);
        // slices of one query share their bound only when there are several
        e.gthr = share && e.S > 1 ? (uint32_t *)(base + O[i].gthr) : nullptr;
        e.cidx = (int32_t *)(base + O[i].cidx);
        e.capx = (float *)(base + O[i].capx);
        e.flags = (int32_t *)(base + O[i].flags);
        e.out_idx = b->oi[i];
        e.out_dist = b->od[i];
    }
    std::vector<int2> tiles(tl.size());
    for (size_t t = 0; t < tl.size(); ++t) tiles[t] = tl[t].t;
    b->ntiles = (int64_t)tiles.size();
    for (int c = 0; c < 3; ++c) b->ncls[c] = (int64_t)tc[c].size();
    b->rows = rows;
    b->queries = queries;
    // merge query lists by private-list size
    std::vector<int32_t> mg[4];
    for (int i = 0; i < ne; ++i) {
        const int cap = merge_cap(E[i]);
        const int c = cap <= 8 ? 0 : (cap <= 16 ? 1 : (cap <= 32 ? 2 : 3));
        for (int64_t q = 0; q < E[i].nq; ++q) mg[c].push_back((int32_t)(E[i].q_base + q));
    }
    std::vector<int32_t> mall;
    for (int c = 0; c < 4; ++c) {
        b->mcount[c] = (int64_t)mg[c].size();
        mall.insert(mall.end(), mg[c].begin(), mg[c].end());
    }
    hipStream_t st = b->cap_stream;
    PN_HIP(hipMemcpyAsync(b->mgids.ensure(std::max<size_t>(1, mall.size()) * 4), mall.data(), mall.size() * 4,
                          hipMemcpyHostToDevice, st));
    PN_HIP(hipMemcpyAsync(b->entries.ensure(ne * sizeof(BatchEntry)), E.data(), ne * sizeof(BatchEntry),
                          hipMemcpyHostToDevice, st));
    PN_HIP(hipMemcpyAsync(b->row_base.ensure(ne * 8), rbase.data(), ne * 8, hipMemcpyHostToDevice, st));
    PN_HIP(hipMemcpyAsync(b->q_base.ensure(ne * 8), qbase.data(), ne * 8, hipMemcpyHostToDevice, st));
    PN_HIP(hipMemcpyAsync(b->tiles.ensure(std::max<size_t>(1, tiles.size()) * 8), tiles.data(), tiles.size() * 8,
                          hipMemcpyHostToDevice, st));
    b->counter.ensure(16); // one queue counter per scan class
    PN_HIP(hipStreamSynchronize(st));
    b->cfg = cfg;

    // capture: no memsets inside (the split kernel clears the tile counter)
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    const BatchEntry *dE = b->entries.as<BatchEntry>();
    const int64_t *dR = b->row_base.as<int64_t>(), *dQ = b->q_base.as<int64_t>();
    int32_t *ctr = b->counter.as<int32_t>();
    const int64_t *nc = b->ncls;
    const unsigned grid_scan =
        (unsigned)std::max<int64_t>(1, std::min<int64_t>(nc[0], (int64_t)b->cus * std::max(1, cfg.wg_per_cu)));
    // the wide scans hold one wave per SIMD (their query fragments in registers): one workgroup
    // per CU
    const unsigned grid_wide = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nc[1], (int64_t)b->cus));
    const unsigned grid_wd = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nc[2], (int64_t)b->cus));
    // The three scans are independent branches of the graph (a fork after the split, a join before
    // the merge): the wide class holds one workgroup -- one wave per SIMD -- per CU for its LDS
    // tiles, and the other two classes' workgroups run beside it instead of after it. The wide
    // scan is launched first so that its workgroups find their LDS free.
    // bit 0: the wide rows by the pipelined scan, bit 1: the narrow rows
    const int pipe = (int)tune_get("batch.pipe");
    const bool fork = tune_get("batch.concurrent_scans") != 0 && nc[1] > 0 && (nc[0] > 0 || nc[2] > 0);
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    struct Cleanup {
        hipStream_t *s;
        hipEvent_t *e;
        ~Cleanup() {
            for (int i = 0; i < 2; ++i)
                if (s[i]) (void)hipStreamDestroy(s[i]);
            for (int i = 0; i < 3; ++i)
                if (e[i]) (void)hipEventDestroy(e[i]);
        }
    } cleanup{side, ev};
    if (fork) {
        for (int i = 0; i < 2; ++i) PN_HIP(hipStreamCreateWithFlags(&side[i], hipStreamNonBlocking));
        for (int i = 0; i < 3; ++i) PN_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
    }
    PN_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(batch_split_kernel, dim3((unsigned)ceil_div(rows, (int64_t)4)), dim3(256), 0, st, dE, dR, ne,
                       rows, ctr);
    const int2 *dt = b->tiles.as<int2>();
    if (fork) {
        PN_HIP(hipEventRecord(ev[0], st));
        PN_HIP(hipStreamWaitEvent(side[0], ev[0], 0));
        PN_HIP(hipStreamWaitEvent(side[1], ev[0], 0));
    }
    if (nc[1] > 0)
        hipLaunchKernelGGL(batch_scan_kernel<1>, dim3(grid_wide), dim3(256), 0, st, dE, dt + nc[0], nc[1], ctr + 1, pipe);
    if (nc[0] > 0)
        hipLaunchKernelGGL(batch_scan_kernel<0>, dim3(grid_scan), dim3(256), 0, fork ? side[0] : st, dE, dt, nc[0],
                           ctr, pipe);
    if (nc[2] > 0)
        hipLaunchKernelGGL(batch_scan_kernel<2>, dim3(grid_wd), dim3(256), 0, fork ? side[1] : st, dE,
                           dt + nc[0] + nc[1], nc[2], ctr + 2, pipe);
    if (fork) {
        PN_HIP(hipEventRecord(ev[1], side[0]));
        PN_HIP(hipEventRecord(ev[2], side[1]));
        PN_HIP(hipStreamWaitEvent(st, ev[1], 0));
        PN_HIP(hipStreamWaitEvent(st, ev[2], 0));
    }
    const bool fused = tune_get("merge.lean") != 0 && tune_get("l2.fuse_rerank") != 0;
    if (fused) {
        hipLaunchKernelGGL(batch_merge_rerank_kernel, dim3(xcd_grid(ceil_div(queries, (int64_t)4))), dim3(256), 0, st,
                           dE, dQ, ne, queries, tune_get("merge.lean") == 2);
    } else if (tune_get("merge.lean") != 0) {
        hipLaunchKernelGGL(batch_merge_lean_kernel, dim3(xcd_grid(ceil_div(queries, (int64_t)4))), dim3(256), 0, st,
                           dE, dQ, ne, queries, tune_get("merge.lean") == 2);
    } else {
        const int32_t *mg = b->mgids.as<int32_t>();
        int64_t o = 0;
        auto merge = [&](auto kc, int c) {
            constexpr int KV = decltype(kc)::value;
            if (b->mcount[c] > 0)
                hipLaunchKernelGGL((batch_merge_kernel<KV>), dim3(xcd_grid(ceil_div(b->mcount[c], (int64_t)4))),
                                   dim3(256), 0, st, dE, dQ, ne, mg + o, b->mcount[c]);
            o += b->mcount[c];
        };
        merge(std::integral_constant<int, 8>{}, 0);
        merge(std::integral_constant<int, 16>{}, 1);
        merge(std::integral_constant<int, 32>{}, 2);
        merge(std::integral_constant<int, 64>{}, 3);
    }
    if (!fused)
        hipLaunchKernelGGL(batch_rerank_kernel, dim3((unsigned)ceil_div(queries, (int64_t)4)), dim3(256), 0, st, dE,
                           dQ, ne, queries);
    hipLaunchKernelGGL(batch_fallback_kernel,
                       dim3((unsigned)std::max<int64_t>(
                           1, std::min<int64_t>(ceil_div(queries, (int64_t)kFallbackChunk), kFallbackMaxGrid))),
                       dim3(256), 0, st, dE,
                       dQ, ne, queries);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(st, &g);
    PN_HIP(ce);
    b->graph = g;
    PN_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
}

static float batch_time(pn_batch *b, int reps) {
    hipEvent_t e0, e1;
    PN_HIP(hipEventCreate(&e0));
    PN_HIP(hipEventCreate(&e1));
    PN_HIP(hipGraphLaunch(b->exec, b->cap_stream)); // warm
    PN_HIP(hipEventRecord(e0, b->cap_stream));
    for (int r = 0; r < reps; ++r) PN_HIP(hipGraphLaunch(b->exec, b->cap_stream));
    PN_HIP(hipEventRecord(e1, b->cap_stream));
    PN_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    PN_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / std::max(1, reps);
}

} // namespace pn

using namespace pn;

extern "C" pn_status pn_batch_create(int32_t n, pn_exact *const *indices, const void *const *queries,
                                     const int64_t *nq, const int32_t *k, int32_t *const *out_idx,
                                     float *const *out_dist, pn_batch **out) {
    return guard([&] {
        PN_REQUIRE(out && n > 0 && indices && queries && nq && k && out_idx && out_dist, PN_EINVAL, "null pointer");
        auto *b = new pn_batch();
        std::unique_ptr<pn_batch> hold(b);
        b->device = indices[0]->device;
        for (int i = 0; i < n; ++i) {
            pn_exact *h = indices[i];
            PN_REQUIRE(h && h->metric == PN_METRIC_L2, PN_EINVAL, "a batch holds exact L2 indices");
            PN_REQUIRE(h->device == b->device, PN_EINVAL, "a batch's indices live on one device");
            PN_REQUIRE(h->n > 0, PN_ERUNTIME, "index must be first initialized with .set() function and non empty dataset");
            PN_REQUIRE(k[i] >= 1 && k[i] <= 32, PN_EINVAL, "batched search serves 1 <= k <= 32");
            PN_REQUIRE(nq[i] >= 1 && queries[i] && out_idx[i] && out_dist[i], PN_EINVAL, "empty or null sub-search");
            b->idx.push_back(h);
            b->gen.push_back(h->generation);
            b->q.push_back((const float *)queries[i]);
            b->nq.push_back(nq[i]);
            b->k.push_back(k[i]);
            b->oi.push_back(out_idx[i]);
            b->od.push_back(out_dist[i]);
        }
        DeviceScope ds(b->device);
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device) == hipSuccess && cus > 0)
            b->cus = cus;
        PN_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
        BatchCfg cfg;
        cfg.slice_rows = (int)tune_get("batch.slice_rows");
        cfg.wg_per_cu = (int)tune_get("batch.wg_per_cu");
        cfg.order = (int)tune_get("batch.order");
        batch_plan(b, cfg);
        *out = hold.release();
    });
}

extern "C" void pn_batch_destroy(pn_batch *b) {
    if (!b) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(b->device);
    delete b;
    if (prev >= 0) (void)hipSetDevice(prev);
}

// The plan's graph holds each index's device rows and sizes as they were at planning time: an
// index set() again since then (new rows, a new allocation, or emptied) makes the plan invalid.
static void require_current(const pn_batch *b) {
    for (size_t i = 0; i < b->idx.size(); ++i)
        PN_REQUIRE(b->idx[i]->generation == b->gen[i], PN_EINVAL,
                   "an index of this batch was set() again after the batch was planned; create a new batch");
}

extern "C" pn_status pn_batch_run(pn_batch *b, void *stream) {
    return guard([&] {
        PN_REQUIRE(b && b->exec, PN_EINVAL, "batch is not planned");
        require_current(b);
        DeviceScope ds(b->device);
        PN_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
    });
}

extern "C" pn_status pn_batch_info(const pn_batch *b, int64_t *tiles, int32_t *cfg /*[3]*/) {
    return guard([&] {
        PN_REQUIRE(b, PN_EINVAL, "batch is null");
        if (tiles) *tiles = b->ntiles;
        if (cfg) {
            cfg[0] = b->cfg.slice_rows;
            cfg[1] = b->cfg.wg_per_cu;
            cfg[2] = b->cfg.order;
        }
    });
}

extern "C" pn_status pn_batch_tune(pn_batch *b, const char *table_path, int32_t reps) {
    return guard([&] {
        PN_REQUIRE(b, PN_EINVAL, "batch is null");
        require_current(b);
        DeviceScope ds(b->device);
        const std::string sig = batch_signature(b);
        if (table_path && *table_path) { // a stored schedule for this shape wins
            std::ifstream f(table_path);
            std::string line;
            while (std::getline(f, line)) {
                std::istringstream ls(line);
                std::string key;
                BatchCfg c;
                if ((ls >> key >> c.slice_rows >> c.wg_per_cu >> c.order) && key == sig) {
                    batch_plan(b, c);
                    return;
                }
            }
        }
        BatchCfg best = b->cfg;
        float best_ms = batch_time(b, std::max(1, (int)reps));
        for (int rows : {128, 256, 512, 1024, 2048, 4096})
            for (int w : {1, 2, 4})
                for (int o : {0, 1, 2}) {
                    BatchCfg c{rows, w, o};
                    batch_plan(b, c);
                    const float ms = batch_time(b, std::max(1, (int)reps));
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = c;
                    }
                }
        batch_plan(b, best);
        if (table_path && *table_path) {
            std::ofstream f(table_path, std::ios::app);
            f << sig << " " << best.slice_rows << " " << best.wg_per_cu << " " << best.order << "\n";
        }
    });
}
