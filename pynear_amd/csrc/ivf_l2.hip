// pn_ivfl2: the search of IVFFlatL2Index (pynear/forest.py:63-115) on the GPU.
//
// The reference loops over queries and probed clusters in Python: a GEMM for the coarse
// distances (:76, :200-204), argpartition to nprobe (:77-80), a GEMV per (query, cluster)
// (:189-194), a per-cluster argpartition and a global heapq (:84-105), ascending squared L2.
// Here one search is four device stages:
//   1. coarse: exact L2 top-nprobe clusters per query (the exact search over the centroids).
//   2. bucketing: (query, probe) pairs are grouped by cluster and cut into work items of at most
//      64 queries of one cluster; items of the longest lists come first.
//   3. scan: per item, a GEMM of the item's queries against the cluster's rows in the form
//      |q|^2 + |x|^2 - 2 q.x on a bf16 copy of the rows (hi only, or hi+lo for 3 products);
//      each (query, probe) keeps its KP best candidates by the approximation.
//   4. rerank: per query the K2 = 64 best candidates over all probes are recomputed in fp32 as
//      sum (q - x)^2 from the fp32 rows (the reference's value, ||q||^2 + ||x||^2 - 2 q.x through
//      BLAS, has no fixed summation order to reproduce) and the k smallest (distance, id) are
//      returned.
// Clusters are stored contiguously in cluster order (ids ascending within a cluster); empty
// clusters are left out, as the reference's per-cluster arrays leave them out (forest.py:51-59).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "common.hpp"
#include "device_util.hpp"
#include "exact_index.hpp"
#include "kernels.hpp"
#include "kmeans_l2.hpp"
#include "l2_mfma.hpp"
#include "ref_arith.hpp"

namespace pn {


// per-lane candidate list of the list scan: 8 for k <= 16, else 16 (a (query, list) pair keeps the
// best of 8 such lists, each over an eighth of the list's rows)
constexpr int kIvfKP = 32; // candidates kept per (query, probe)
constexpr int kIvfK2 = 64; // candidates reranked per query

} // namespace pn

struct pn_ivfl2 {
    int device = 0;
    int cus = 256;
    int64_t n = 0, d = 0;
    int32_t nl = 0;          // non-empty clusters
    pn_exact cent;           // coarse index over the non-empty clusters' centroids
    pn::DevBuf X;            // fp32 rows, original order (rerank)
    pn::L2Set lrows;         // bf16 hi/lo of the rows in cluster order + norms
    pn::DevBuf lids, loff;   // cluster-order position -> original id; cluster row offsets [nl + 1]
    pn::DevBuf corder;       // cluster ids by decreasing list length (work-item order)
    pn::DevBuf clist;        // k-means cluster id -> list index (-1: empty cluster), nc entries
    int32_t nc = 0;          // k-means clusters (empty ones included)
    pn::DevBuf pmap;         // per search: given probes mapped to list indices
    pn::DevBuf l8;           // fp8 e4m3 residuals x - c (cluster order) scaled by sx8, dp bytes per row
    // segments (row ranges: a list's parts, or the list): fp8 centre per segment (l8's layout, one
    // row per segment), {R, X, |centre|, min |x - c|^2}, row offsets [nseg + 1], first segment per list
    pn::DevBuf c8, smeta, soff, sbase;
    int64_t nseg = 0;
    pn::DevBuf rn8;          // |x - c|^2 per row (cluster order)
    pn::DevBuf lcent;        // list centroids, fp32 nl x d (list order)
    float sx8 = 1.f;
    pn::DevBuf qpad, qsplit_h, qsplit_l, qn, probes, pdist, lcount, cursor, pair_q, pair_p, pair_c, pslot;
    pn::DevBuf items, ntot /* item ranges */, itoff, woff, cand, qoff, qthr, qres8, qresn, qkey, qorder, ohist, oidx,
        odist, ocnt, qres8n, pact;
    bool ohist_dirty = false; // the rerank order's histogram may be non-zero (a search stopped early)
    pn::DevBuf work;         // the last fp8 scan launch: per workgroup MFMA instructions and fetched row bytes
    int64_t work_n = 0;      // its workgroups
};

namespace pn {

// Pairs are bucketed by cluster in chunks of kPairChunk per workgroup: with at most kLdsBins
// clusters the counts of a chunk are gathered in LDS first, so a cluster that many queries probe
// takes one global atomic per chunk instead of one per pair.
constexpr int kPairChunk = 1024;
constexpr int kLdsBins = 8192;

// The pairs of a pass are the probe slots p_lo .. p_lo + npw - 1 of every query (rows of np
// slots): pass pair t is slot p_lo + t % npw of query t / npw.
__device__ __forceinline__ int64_t pass_slot(int64_t t, int np, int p_lo, int npw) {
    return npw == np ? t : (t / npw) * np + p_lo + t % npw;
}

__global__ __launch_bounds__(256) void probe_count_kernel(const int32_t *__restrict__ probes, int64_t total, int np,
                                                          int p_lo, int npw, int nl, int32_t *__restrict__ lcount) {
    __shared__ int32_t hist[kLdsBins];
    const int64_t c0 = (int64_t)blockIdx.x * kPairChunk, c1 = min(total, c0 + kPairChunk);
    if (nl > kLdsBins) {
        for (int64_t t = c0 + threadIdx.x; t < c1; t += 256) {
            const int pc = probes[pass_slot(t, np, p_lo, npw)];
            if (pc >= 0) atomicAdd(&lcount[pc], 1);
        }
        return;
    }
    for (int b = threadIdx.x; b < nl; b += 256) hist[b] = 0;
    __syncthreads();
    for (int64_t t = c0 + threadIdx.x; t < c1; t += 256) {
        const int pc = probes[pass_slot(t, np, p_lo, npw)];
        if (pc >= 0) atomicAdd(&hist[pc], 1); // -1: no probe in this slot
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nl; b += 256)
        if (hist[b]) atomicAdd(&lcount[b], hist[b]);
}

constexpr int kIvfIQ = 64;   // queries (pairs of one cluster) per work item of the bf16 scan
constexpr int kIvf8IQ = 128; // ... of the fp8 scan
constexpr int kIvf8Cap = 64; // candidates kept per (query, probe) by the fp8 scan
constexpr int kIvfKC = 64;  // dims per K chunk
constexpr int kIvfQS = 72;  // LDS row stride (bf16) of a staged query chunk: 16 B of pad spreads the banks

// Work items are (cluster, first pair, count <= iq), cluster by cluster in the order corder
// (lists by decreasing length, so the longest items come first). build_items_kernel (one
// workgroup) scans the per-cluster pair, item and work counts (an item's work is its list length
// + 64) and sets each cluster's first pair slot (pair cursor); expand_items_kernel writes the items
// in parallel. Items are split into nranges contiguous ranges of about equal work. With
// nranges = kXcds, range b is run by the workgroups with id % kXcds == b, which round-robin
// dispatch places on one XCD, so a long list that many queries probe is read into one or two L2s
// instead of all eight; with nranges = 1 (the default: measured faster, the per-XCD ranges
// balance worse than one queue) every workgroup takes from one range. bins holds per range its
// first item, its item count and a zeroed work counter; bins[3 kXcds] the item total,
// bins[3 kXcds + 1] the pair total and bins[3 kXcds + 2 ..] (as int64) the total work.
constexpr int kXcds = 8;
constexpr int kBinsInts = 3 * kXcds + 4;
__global__ __launch_bounds__(256) void build_items_kernel(const int32_t *__restrict__ lcount,
                                                          const int32_t *__restrict__ corder,
                                                          const int64_t *__restrict__ loff, int nl,
                                                          int32_t *__restrict__ cursor, int32_t *__restrict__ itoff,
                                                          int64_t *__restrict__ woff, int32_t *__restrict__ bins,
                                                          int nranges, int iq, int32_t *__restrict__ ohist, int nh) {
    __shared__ int32_t s_pairs[256], s_items[256];
    __shared__ int64_t s_work[256];
    const int tid = threadIdx.x;
    const int per = (nl + 255) / 256;
    const int o0 = min(nl, tid * per), o1 = min(nl, o0 + per);
    int pairs = 0, nit = 0;
    int64_t work = 0;
    for (int o = o0; o < o1; ++o) {
        const int c = corder[o], m = lcount[c];
        const int it = (m + iq - 1) / iq;
        pairs += m;
        nit += it;
        work += (int64_t)it * (loff[c + 1] - loff[c] + 64);
    }
    s_pairs[tid] = pairs;
    s_items[tid] = nit;
    s_work[tid] = work;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) { // inclusive scans
        const int a = tid >= off ? s_pairs[tid - off] : 0, b = tid >= off ? s_items[tid - off] : 0;
        const int64_t w = tid >= off ? s_work[tid - off] : 0;
        __syncthreads();
        s_pairs[tid] += a;
        s_items[tid] += b;
        s_work[tid] += w;
        __syncthreads();
    }
    int pa = s_pairs[tid] - pairs, ia = s_items[tid] - nit;
    int64_t wa = s_work[tid] - work;
    for (int o = o0; o < o1; ++o) {
        const int c = corder[o], m = lcount[c];
        const int it = (m + iq - 1) / iq;
        cursor[c] = pa;
        itoff[o] = ia;
        woff[o] = wa;
        pa += m;
        ia += it;
        wa += (int64_t)it * (loff[c + 1] - loff[c] + 64);
    }
    if (tid < kXcds) { // one thread per range writes its three words
        const bool all = nranges == 1 && tid == 0; // a single range holds every item
        bins[tid] = all ? 0 : INT32_MAX;              // first item (else lowered by expand_items_kernel)
        bins[kXcds + tid] = all ? s_items[255] : 0;   // items in the range
        bins[2 * kXcds + tid] = 0;                    // next item to take
    }
    if (tid == 255) {
        bins[3 * kXcds] = s_items[255];
        bins[3 * kXcds + 1] = s_pairs[255];
        *(int64_t *)(bins + 3 * kXcds + 2) = s_work[255];
    }
    if (ohist) { // the rerank's query order: exclusive scan of the first-probe histogram, in place
        __syncthreads();
        const int hp = (nh + 255) / 256, h0 = min(nh, tid * hp), h1 = min(nh, h0 + hp);
        int sum = 0;
        for (int o = h0; o < h1; ++o) sum += ohist[o];
        s_pairs[tid] = sum;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const int a = tid >= off ? s_pairs[tid - off] : 0;
            __syncthreads();
            s_pairs[tid] += a;
            __syncthreads();
        }
        int run = s_pairs[tid] - sum;
        for (int o = h0; o < h1; ++o) {
            const int v = ohist[o];
            ohist[o] = run;
            run += v;
        }
    }
}

// One thread per item g: its cluster (the last corder position o with itoff[o] <= g), its pairs
// and its range.
__global__ void expand_items_kernel(const int32_t *__restrict__ lcount, const int32_t *__restrict__ corder,
                                    const int64_t *__restrict__ loff, int nl, const int32_t *__restrict__ cursor,
                                    const int32_t *__restrict__ itoff, const int64_t *__restrict__ woff,
                                    int64_t max_items, int nranges, int iq, int4 *__restrict__ items,
                                    int32_t *__restrict__ bins, int eblocks, const int32_t *__restrict__ qkey,
                                    int64_t nq, int32_t *__restrict__ ocur, int32_t *__restrict__ qorder) {
    if ((int)blockIdx.x >= eblocks) { // the rerank's query order: queries by first probe
        const int64_t q = (int64_t)(blockIdx.x - eblocks) * blockDim.x + threadIdx.x;
        if (q < nq) {
            const int at = atomicAdd(&ocur[qkey[q]], 1);
            if (at < nq) qorder[at] = (int32_t)q;
        }
        return;
    }
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= max_items || g >= bins[3 * kXcds]) return;
    int lo = 0, hi = nl - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (itoff[mid] <= g) lo = mid; else hi = mid - 1;
    }
    const int c = corder[lo], m = lcount[c];
    const int s = (int)(g - itoff[lo]);
    items[g] = make_int4(c, cursor[c] + iq * s, min(iq, m - iq * s), 0);
    if (nranges > 1) { // (with one range, build_items_kernel has set it to all items)
        const int64_t W = max<int64_t>(*(const int64_t *)(bins + 3 * kXcds + 2), 1);
        const int64_t wi = loff[c + 1] - loff[c] + 64;
        const int b = min(nranges - 1, (int)((double)(woff[lo] + s * wi + wi / 2) * nranges / (double)W));
        atomicMin(&bins[b], (int)g); // ranges are contiguous: b is monotone in g
        atomicAdd(&bins[kXcds + b], 1);
    }
}

// Writes each pair (query, probe rank, and its list) to its cluster's next free slot, and the
// slot into pslot[query np + probe rank]. Per chunk, one global
// atomic per cluster reserves the chunk's slots (LDS path); pairs then take ranks inside them.
__global__ __launch_bounds__(256) void probe_scatter_kernel(const int32_t *__restrict__ probes, int64_t total, int np,
                                                            int p_lo, int npw, int nl, int32_t *__restrict__ cursor,
                                                            int32_t *__restrict__ pair_q, int32_t *__restrict__ pair_p,
                                                            int32_t *__restrict__ pair_c, int32_t *__restrict__ pslot,
                                                            int32_t *__restrict__ lcount) {
    __shared__ int32_t hist[kLdsBins];
    // the list counts' last reader (expand_items_kernel) has run: leave them zero for the next pass
    for (int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x; b < nl; b += (int64_t)gridDim.x * 256) lcount[b] = 0;
    const int64_t c0 = (int64_t)blockIdx.x * kPairChunk, c1 = min(total, c0 + kPairChunk);
    if (nl > kLdsBins) {
        for (int64_t t = c0 + threadIdx.x; t < c1; t += 256) {
            const int64_t sl = pass_slot(t, np, p_lo, npw);
            if (probes[sl] < 0) continue;
            const int pos = atomicAdd(&cursor[probes[sl]], 1);
            pair_q[pos] = (int32_t)(sl / np);
            pair_p[pos] = (int32_t)(sl % np);
            pair_c[pos] = probes[sl];
            pslot[sl] = pos;
        }
        return;
    }
    for (int b = threadIdx.x; b < nl; b += 256) hist[b] = 0;
    __syncthreads();
    for (int64_t t = c0 + threadIdx.x; t < c1; t += 256) {
        const int pc = probes[pass_slot(t, np, p_lo, npw)];
        if (pc >= 0) atomicAdd(&hist[pc], 1);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nl; b += 256) // count -> first global slot of this chunk's run
        if (hist[b]) hist[b] = atomicAdd(&cursor[b], hist[b]);
    __syncthreads();
    for (int64_t t = c0 + threadIdx.x; t < c1; t += 256) {
        const int64_t sl = pass_slot(t, np, p_lo, npw);
        if (probes[sl] < 0) continue;
        const int pos = atomicAdd(&hist[probes[sl]], 1);
        pair_q[pos] = (int32_t)(sl / np);
        pair_p[pos] = (int32_t)(sl % np);
        pair_c[pos] = probes[sl];
        pslot[sl] = pos;
    }
}

// Candidate slot of (query q, probe p): q np + p for np probes per query, or qoff[q] + p when
// queries have their own counts (qoff: exclusive prefix of the per-query probe spans).
__device__ __forceinline__ int64_t cand_slot(const int32_t *__restrict__ qoff, int64_t q, int np, int p) {
    return qoff ? (int64_t)qoff[q] + p : q * np + p;
}

// Per-query probe span (last used slot + 1; 0 for none) into span[q], span[nq] = 0; an exclusive
// scan of span then gives qoff.
__global__ void probe_span_kernel(const int32_t *__restrict__ probes, int64_t nq, int np, int32_t *__restrict__ span) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q > nq) return;
    int n = 0;
    if (q < nq)
        for (int j = np - 1; j >= 0; --j)
            if (probes[q * np + j] >= 0) {
                n = j + 1;
                break;
            }
    span[q] = n;
}

// Fills the candidate slots of every pair (qoff[nq] of them, cap entries each) with v.
__global__ void fill_slots_kernel(uint64_t *__restrict__ p, const int32_t *__restrict__ qoff, int64_t nq, int cap,
                                  uint64_t v) {
    const int64_t n = (int64_t)qoff[nq] * cap;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Bound for the second pass: per query, the k-th best approximation among the cap candidate
// slots of its first probe (k distinct rows are at least that close) times a margin for the
// approximation's error; a row of another list whose approximation is above it cannot be among
// the k nearest. kKey32None when that probe has fewer than k candidates. Keys below the
// returned value are kept. One wave per query; cap <= 256.
__global__ __launch_bounds__(256) void pass_bound_kernel(const uint64_t *__restrict__ cand, int64_t nq, int np, int cap,
                                                         const int32_t *__restrict__ qoff, int k, float margin,
                                                         uint32_t *__restrict__ qthr) {
    const int lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= nq) return;
    const uint64_t *src = cand + cand_slot(qoff, q, np, 0) * cap;
    uint64_t L[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) L[j] = kKey64None;
    for (int e = lane; e < cap; e += 64) list_insert_u64<4>(L, src[e]);
    uint64_t m = kKey64None;
    for (int r = 0; r < k; ++r) {
        m = L[0];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const uint64_t o = __shfl_xor(m, off, 64);
            m = o < m ? o : m;
        }
        if (m == kKey64None) break;
        if (L[0] == m) {
#pragma unroll
            for (int j = 0; j < 3; ++j) L[j] = L[j + 1];
            L[3] = kKey64None;
        }
    }
    if (lane != 0) return;
    if (m == kKey64None) {
        qthr[q] = kKey32None;
        return;
    }
    const float v = key_f32((uint32_t)(m >> 32) | 31u) * margin; // (the key's largest value)
    qthr[q] = isfinite(v) ? (f32_key(v) | 31u) + 1u : kKey32None;
}

// One wave per query: the bound of a later pass is the K2-th smallest candidate key of the
// query's first probe (cap <= 64 slots, sorted or not). At least K2 candidates of the first probe
// are at or below it, so no candidate above it is among the K2 the rerank takes; the scans keep
// keys below qthr = its 32-bit key + 1. kKey32None when the first probe has fewer than K2.
__global__ __launch_bounds__(256) void pass_bound_k2_kernel(const uint64_t *__restrict__ cand, int64_t nq, int np,
                                                            int cap, const int32_t *__restrict__ qoff, int K2,
                                                            uint32_t *__restrict__ qthr) {
    const int lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= nq) return;
    const uint64_t *src = cand + cand_slot(qoff, q, np, 0) * cap;
    uint64_t v[1] = {lane < cap ? src[lane] : kKey64None};
    wave_sort_u64<1, 2>(v, lane);
    const uint64_t m = __shfl(v[0], K2 - 1, 64);
    if (lane == 0) qthr[q] = m == kKey64None ? kKey32None : (uint32_t)(m >> 32) + 1u;
}

// The next work item of a workgroup of range xb: from its own range while that has items, then
// from the other ranges in turn (a range's counter may run past its count). -1: none left.
__device__ __forceinline__ int take_item(int32_t *__restrict__ bins, int nranges, int xb) {
    for (int r = 0; r < nranges; ++r) {
        const int b = xb + r < nranges ? xb + r : xb + r - nranges;
        const int count = bins[kXcds + b];
        if (count == 0) continue;
        const int slot = atomicAdd(&bins[2 * kXcds + b], 1);
        if (slot < count) return bins[b] + slot;
    }
    return -1;
}


// Candidate slots of (query, probe) pairs without a probe stay empty.
__global__ void fill_u64_kernel(uint64_t *__restrict__ p, int64_t n, uint64_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Given probes (cluster ids, counts[q] valid per row of np_max) -> list indices, -1 for none.
__global__ void map_probes_kernel(const int32_t *__restrict__ probes, const int32_t *__restrict__ counts,
                                  int64_t nq, int np_max, const int32_t *__restrict__ clist, int nc,
                                  int32_t *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq * np_max) return;
    const int64_t q = t / np_max;
    const int j = (int)(t - q * np_max), c = probes[t];
    out[t] = j < counts[q] && c >= 0 && c < nc ? clist[c] : -1;
}

// Inserts one tile's 16 keys of one query into the lane's sorted list (Lk, Li) of KL. A key is
// the approximation's float bits with the low 5 bits replaced by the row's offset in the tile, so
// a minimum names its row (base + (key & 31)). Keys at or above thr are dropped; thr may be below
// the list's last key when it comes from another list of the same query (the lists together
// then still hold KL keys below it). The loop runs while any lane of the wave has a key below
// its thr. FIRST: the list is empty and takes the tile's keys through a sorting network.
template <int KL, bool FIRST>
__device__ __forceinline__ void list_insert16(uint32_t (&key)[16], uint32_t (&Lk)[KL], uint32_t (&Li)[KL],
                                              uint32_t base, uint32_t &thr) {
    if (FIRST) {
        sort16(key);
#pragma unroll
        for (int j = 0; j < KL; ++j) {
            Lk[j] = j < 16 && key[j] < thr ? key[j] : kKey32None;
            Li[j] = j < 16 ? base + (key[j] & 31u) : 0u;
        }
        thr = min(thr, Lk[KL - 1]);
        return;
    }
    uint32_t mn = kKey32None;
#pragma unroll
    for (int r = 0; r < 16; ++r) mn = min(mn, key[r]);
    while (__any(mn < thr)) {
        if (mn < thr) {
#pragma unroll
            for (int r = 0; r < 16; ++r) key[r] = key[r] == mn ? kKey32None : key[r];
            const uint32_t ir = base + (mn & 31u);
#pragma unroll
            for (int j = KL - 1; j > 0; --j) {
                const bool up = mn < Lk[j - 1];
                const bool here = mn < Lk[j];
                Li[j] = up ? Li[j - 1] : (here ? ir : Li[j]);
                Lk[j] = max(Lk[j - 1], min(mn, Lk[j]));
            }
            if (mn < Lk[0]) {
                Li[0] = ir;
                Lk[0] = mn;
            }
            thr = min(thr, Lk[KL - 1]);
        }
        mn = kKey32None;
#pragma unroll
        for (int r = 0; r < 16; ++r) mn = min(mn, key[r]);
    }
}

// Row block done: keys |q|^2 + |x|^2 - 2 q.x (q.x = acc * dscale of the query) of the wave's
// 64 rows (from row rbase; rows at or past r1 are masked) for its two queries per lane, into the
// lane's lists. Thresholds are shared by the 8 lists of a query: the other half-wave's at once,
// the other waves' as posted after their previous row block (and a query may start with a bound
// from elsewhere). A (tile, query) whose smallest approximation is at or above every lane's
// threshold is done after that test: most tiles are, once the lists have filled. TWO = false:
// the item has at most 32 queries and the second query tile is not computed.
// s_thr[l * tstride + q]: the threshold of list l (of nlist per query) of query q; the lane's
// lists are number `list` of queries qcol and qcol + 32. norm_lds: the row norms are read from
// snorm (LDS, the list's rows from r0) instead of lxn (global memory).
template <int KL>
__device__ __forceinline__ void rowblock_keys(const f32x16 (&acc)[2][2], float dscale0, float dscale1,
                                              const float *__restrict__ lxn, int64_t rbase, int64_t r1, bool first,
                                              float qn0, float qn1, uint32_t (&Lk)[2][KL], uint32_t (&Li)[2][KL],
                                              uint32_t *s_thr, int tstride, int nlist, int list, int qcol, int half,
                                              bool two, const float *snorm = nullptr, bool norm_lds = false,
                                              int64_t r0 = 0) {
    const int rlim = (int)min<int64_t>(r1 - rbase, 64); // rows of this wave's block inside the list
    uint32_t thr[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        uint32_t m = Lk[j][KL - 1];
        for (int l = 0; l < nlist; ++l) m = min(m, s_thr[l * tstride + qcol + 32 * j]);
        thr[j] = m;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        // 32-bit offsets from the tile's first row; rows past the list end are clamped / masked
        const int lim = rlim - 32 * i; // rows of this tile inside the list (may be <= 0)
        if (lim <= 0) break;           // (wave-uniform; so is every later tile of the block)
        float xr[16];
        if (norm_lds) { // the list's norms staged in LDS from row r0 (LDS loads, not flat ones)
            typedef const __attribute__((address_space(3))) float lds_float;
            lds_float *xs = (lds_float *)snorm + (int)(rbase - r0) + 32 * i;
#pragma unroll
            for (int r = 0; r < 16; ++r) xr[r] = xs[min(8 * (r >> 2) + 4 * half + (r & 3), lim - 1)];
        } else {
            const float *xb = lxn + rbase + 32 * i;
#pragma unroll
            for (int r = 0; r < 16; ++r) xr[r] = xb[min(8 * (r >> 2) + 4 * half + (r & 3), lim - 1)];
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j == 1 && !two) break;
            const float qv = j == 0 ? qn0 : qn1;
            const float m2 = -2.f * (j == 0 ? dscale0 : dscale1);
            float v[16];
            float vmin = INFINITY;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                v[r] = fmaxf(0.f, fmaf(m2, acc[i][j][r], qv + xr[r]));
                vmin = fminf(vmin, v[r]);
            }
            thr[j] = min(thr[j], (uint32_t)__shfl_xor((int)thr[j], 32, 64));
            // every key of the tile is >= bits(vmin) with the offset bits cleared (masked rows
            // only lower vmin)
            if (!__any((f32_key(vmin) & ~31u) < thr[j])) continue;
            uint32_t key[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) key[r] = (f32_key(v[r]) & ~31u) | (uint32_t)(8 * (r >> 2) + 4 * half + (r & 3));
            if (lim < 32)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    key[r] = 8 * (r >> 2) + 4 * half + (r & 3) < lim ? key[r] : kKey32None;
            const uint32_t base = (uint32_t)(rbase + 32 * i);
            if (first && i == 0)
                list_insert16<KL, true>(key, Lk[j], Li[j], base, thr[j]);
            else
                list_insert16<KL, false>(key, Lk[j], Li[j], base, thr[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) s_thr[list * tstride + qcol + 32 * j] = thr[j];
}

// Item done: the lists of each query (nlist per query: waves x 2 halves) go to LDS (smem; the
// query buffers are no longer read), at sl[(l * tstride + q) * KL], then one lane per query
// merges its kp best into cand[query][probe] (kp slots per pair).
template <int KL, int NLMAX>
__device__ __forceinline__ void merge_item_lists(unsigned char *smem, const uint32_t (&Lk)[2][KL],
                                                 const uint32_t (&Li)[2][KL], const int32_t *s_q,
                                                 const int32_t *__restrict__ pair_p, int cnt, int np,
                                                 const int32_t *__restrict__ qoff, uint64_t *__restrict__ cand,
                                                 int tstride, int nlist, int list, int qcol, int kp) {
    const int tid = threadIdx.x;
    uint64_t *sl = (uint64_t *)smem;
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int e = 0; e < KL; ++e)
            sl[(list * tstride + qcol + 32 * j) * KL + e] =
                Lk[j][e] == kKey32None ? kKey64None : (((uint64_t)Lk[j][e] << 32) | Li[j][e]);
    __syncthreads();
    if (tid < cnt) {
        uint64_t *out = cand + cand_slot(qoff, s_q[tid], np, pair_p[tid]) * kp;
        int p[NLMAX];
        uint64_t head[NLMAX];
#pragma unroll
        for (int l = 0; l < NLMAX; ++l) {
            p[l] = 0;
            head[l] = l < nlist ? sl[(l * tstride + tid) * KL] : kKey64None;
        }
        for (int e = 0; e < kp; ++e) {
            int best = 0;
            uint64_t bv = head[0];
#pragma unroll
            for (int l = 1; l < NLMAX; ++l)
                if (head[l] < bv) {
                    bv = head[l];
                    best = l;
                }
            out[e] = bv;
#pragma unroll
            for (int l = 0; l < NLMAX; ++l)
                if (l == best && bv != kKey64None) {
                    p[l] += 1;
                    head[l] = p[l] < KL ? sl[(l * tstride + tid) * KL + p[l]] : kKey64None;
                }
        }
    }
}

typedef int v8i_t __attribute__((ext_vector_type(8)));
union V8 {
    v8i_t v;
    uint4 u[2];
};

// Row block done in the fp8 scan: keys of the wave's two 32-row tiles (rows tb0 .. tb0 + 63, tb0
// a multiple of 32; rows outside [r0, r1) are masked, tiles at or past r1 skipped) for its two
// query tiles, into the lane's lists. Lane (col, half) holds query column qcol of tile j and rows
// 32 i + 8 (r / 4) + 4 half + r % 4 of element r. v = max(0, fma(m2, acc, xr) + qv) approximates
// |q - x|^2 (m2 = -2 / scale^2, xr = |x - c|^2 from xs[row - tb0], qv = |q - c|^2). A (tile,
// query) is done after one test when max(0, min_r fma(m2, acc_r, xr_r) + qv), a lower bound of
// its 16 values whatever rows are masked, is at or above the threshold: most tiles, once the
// lists hold good keys or under a bound from an earlier pass. A key is v's bits with the low 5
// bits replaced by the row's offset in its tile. Thresholds: the lane's own, its half-wave
// partner's and the smallest any list of the query posted in sthr[query] (LDS), where the lane
// posts its own. FIRST: the lists are empty.
//
// Keys here carry the element index r (0 .. 15, an inline constant) in their low 4 bits instead
// of the row offset (so no per-element offsets live in registers); the row, tb + 4 half +
// 8 (r / 4) + r % 4, is formed only when a key enters a list (list_insert16r).
template <int KL, bool FIRST>
__device__ __forceinline__ void list_insert16r(uint32_t (&key)[16], uint32_t (&Lk)[KL], uint32_t (&Li)[KL],
                                               uint32_t base_h, uint32_t &thr) {
    auto row_of = [&](uint32_t k) { return base_h + (k & 3u) + ((k & 12u) << 1); };
    if (FIRST) {
        sort16(key);
#pragma unroll
        for (int j = 0; j < KL; ++j) {
            Lk[j] = j < 16 && key[j] < thr ? key[j] : kKey32None;
            Li[j] = j < 16 ? row_of(key[j]) : 0u;
        }
        thr = min(thr, Lk[KL - 1]);
        return;
    }
    uint32_t mn = kKey32None;
#pragma unroll
    for (int r = 0; r < 16; ++r) mn = min(mn, key[r]);
    while (__any(mn < thr)) {
        if (mn < thr) {
#pragma unroll
            for (int r = 0; r < 16; ++r) key[r] = key[r] == mn ? kKey32None : key[r];
            const uint32_t ir = row_of(mn);
#pragma unroll
            for (int j = KL - 1; j > 0; --j) {
                const bool up = mn < Lk[j - 1];
                const bool here = mn < Lk[j];
                Li[j] = up ? Li[j - 1] : (here ? ir : Li[j]);
                Lk[j] = max(Lk[j - 1], min(mn, Lk[j]));
            }
            if (mn < Lk[0]) {
                Li[0] = ir;
                Lk[0] = mn;
            }
            thr = min(thr, Lk[KL - 1]);
        }
        mn = kKey32None;
#pragma unroll
        for (int r = 0; r < 16; ++r) mn = min(mn, key[r]);
    }
}

// Row norms come from LDS only: a global load here would enter the vector-memory counter behind
// the fragment loads in flight, and waiting for it would drain them at every row block. Rows are
// 32-bit (a list's rows and the database's fit in int32).
template <int KL>
__device__ __forceinline__ void fp8_rowblock_keys(const f32x16 (&acc)[2][2], float m2, const float *xs, int tb0,
                                                  int r0, int r1, bool first, float qv0, float qv1,
                                                  uint32_t (&Lk)[2][KL], uint32_t (&Li)[2][KL], uint32_t *sthr,
                                                  int qcol, int half, bool two) {
    uint32_t thr[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) thr[j] = min(Lk[j][KL - 1], sthr[qcol + 32 * j]);
    const int h4 = 4 * half;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int tb = tb0 + 32 * i;
        if (tb >= r1) break; // (wave-uniform; so is the next tile)
        // the tile's rows inside the list are offsets lo .. hi - 1 (all 32 except at a list's ends)
        const int lo = max(r0 - tb, 0), hi = min(r1 - tb, 32);
        const bool edge = lo > 0 || hi < 32;
        float xr[16];
        {
            typedef const __attribute__((address_space(3))) float lds_float;
            lds_float *x = (lds_float *)xs + 32 * i + 4 * half;
#pragma unroll
            for (int r = 0; r < 16; ++r) xr[r] = x[8 * (r >> 2) + (r & 3)];
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j == 1 && !two) break;
            const float qv = j == 0 ? qv0 : qv1;
            float u[16];
            float umin = INFINITY;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                u[r] = fmaf(m2, acc[i][j][r], xr[r]);
                umin = fminf(umin, u[r]);
            }
            thr[j] = min(thr[j], (uint32_t)__shfl_xor((int)thr[j], 32, 64));
            if (!__any((f32_key(fmaxf(0.f, umin + qv)) & ~15u) < thr[j])) continue;
            uint32_t key[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) key[r] = (f32_key(fmaxf(0.f, u[r] + qv)) & ~15u) | (uint32_t)r;
            if (edge) { // (uniform; a list's first and last tiles)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int off = h4 + 8 * (r >> 2) + (r & 3);
                    key[r] = off >= lo && off < hi ? key[r] : kKey32None;
                }
            }
            if (first && i == 0)
                list_insert16r<KL, true>(key, Lk[j], Li[j], (uint32_t)tb + h4, thr[j]);
            else
                list_insert16r<KL, false>(key, Lk[j], Li[j], (uint32_t)tb + h4, thr[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j)
        if (j == 0 || two) atomicMin(&sthr[qcol + 32 * j], thr[j]);
}

// fp8 rows are stored by 32-row groups and 64-dim chunks: group g, chunk ch is one 2 KB block of
// its 32 rows' 64 bytes, so a wave's fragment loads of one chunk read whole blocks.
__device__ __forceinline__ int64_t fp8_row_chunk(int64_t row, int ch, int nch) {
    return (((row >> 5) * nch + ch) << 11) + ((row & 31) << 6);
}

// Rows as residuals to their list's centroid, r = x - c (cluster order; the row's list is found in
// loff). MODE 0: the largest |r| over all rows into *amax (float bits). MODE 1: |r|^2 into rn[row]
// and r * scale as fp8 e4m3 (OCP, round to nearest even, clamped to +-448) into L8 (dp bytes per
// row with zero padding, laid out by fp8_row_chunk). One wave per row.
template <int MODE>
__global__ __launch_bounds__(256) void residual_kernel(const float *__restrict__ X, int64_t n, int d, int dp,
                                                       const int64_t *__restrict__ loff, int nl,
                                                       const float *__restrict__ lcent, float scale,
                                                       unsigned int *__restrict__ amax, uint8_t *__restrict__ L8,
                                                       float *__restrict__ rn) {
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    int lo = 0, hi = nl - 1; // last list with loff[list] <= row
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (loff[mid] <= row) lo = mid; else hi = mid - 1;
    }
    const float *x = X + row * d, *c = lcent + (int64_t)lo * d;
    if (MODE == 0) {
        float m = 0.f;
        for (int e = lane; e < d; e += 64) m = fmaxf(m, fabsf(x[e] - c[e]));
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
        if (lane == 0) atomicMax(amax, __float_as_uint(m));
        return;
    }
    float ss = 0.f;
    for (int e = 4 * lane; e < dp; e += 256) {
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float r = e + u < d ? x[e + u] - c[e + u] : 0.f;
            ss = fmaf(r, r, ss);
            v[u] = fminf(fmaxf(r * scale, -448.f), 448.f);
        }
        int w = 0;
        w = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], w, false);
        w = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], w, true);
        *(int *)(L8 + fp8_row_chunk(row, e / kIvfKC, dp / kIvfKC) + e % kIvfKC) = w;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off, 64);
    if (lane == 0) rn[row] = ss;
}

// Query residuals of every pair, (q - c) * scale as fp8 e4m3 (clamped to +-448) into R8[pair
// slot] (dp bytes, zero padding), |q - c|^2 into Rn[pair slot] and, when Rq is given, the norm of
// the fp8 residual (its fp32 sum rounded up by 2^-10) into Rq[pair slot]. One wave per query: its row is
// read once (dims 4 l + 256 t .. +3 in lane l), then each of its probes p_lo .. p_lo + npw - 1
// (list probes[q np + p], slot pslot[q np + p]) is converted; slots of absent probes are skipped.
// FULL: d == dp == 256 NV with 16-byte aligned rows, so no dimension is masked.
template <int NV, bool FULL>
__global__ __launch_bounds__(256) void query_residual_kernel(const float *__restrict__ Q, int64_t qstride, int d,
                                                             int dp, const float *__restrict__ lcent, float scale,
                                                             const int32_t *__restrict__ probes, int64_t nq, int np,
                                                             int p_lo, int npw, const int32_t *__restrict__ pslot,
                                                             uint8_t *__restrict__ R8, float *__restrict__ Rn,
                                                             float *__restrict__ Rq, bool xsplit) {
    // xsplit: eight workgroups per four queries, workgroup b (on XCD b % 8) converting only the
    // probes of lists c with c % 8 == b % 8, so each XCD's L2 holds an eighth of the centroids
    // and a query of many probes reads its row once per XCD instead of a centroid row per probe
    const int64_t q = (int64_t)(xsplit ? blockIdx.x >> 3 : blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int xcd = (int)(blockIdx.x & 7);
    const int lane = threadIdx.x & 63;
    if (q >= nq) return;
    const float *qrow = Q + q * qstride;
    const bool vec = FULL || ((d & 3) == 0 && (qstride & 3) == 0); // 16-byte loads (rows of whole float4)
    float qv[NV][4]; // NV = dp / 256 rounded up
#pragma unroll
    for (int t = 0; t < NV; ++t) {
        const int e0 = 4 * lane + 256 * t;
        if (vec) {
            const float4 v = FULL || e0 < d ? *(const float4 *)(qrow + e0) : make_float4(0.f, 0.f, 0.f, 0.f);
            qv[t][0] = v.x, qv[t][1] = v.y, qv[t][2] = v.z, qv[t][3] = v.w;
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) qv[t][u] = e0 + u < d ? qrow[e0 + u] : 0.f;
        }
    }
    // probes in chunks of 64: lane j reads probe j's cluster and slot once (no dependent loads per
    // probe), and the next probe's centroid row is loaded while this one's residual is made (two
    // register sets rotated by name)
    typedef float Row[NV][4];
    auto load_c = [&](int c, Row &cv) __attribute__((always_inline)) {
        const float *cc = lcent + (int64_t)max(c, 0) * d;
#pragma unroll
        for (int t = 0; t < NV; ++t) {
            const int e0 = 4 * lane + 256 * t;
            if (vec) {
                const float4 c4 = c >= 0 && (FULL || e0 < d) ? *(const float4 *)(cc + e0) : make_float4(0.f, 0.f, 0.f, 0.f);
                cv[t][0] = c4.x, cv[t][1] = c4.y, cv[t][2] = c4.z, cv[t][3] = c4.w;
            } else {
#pragma unroll
                for (int u = 0; u < 4; ++u) cv[t][u] = c >= 0 && e0 + u < d ? cc[e0 + u] : 0.f;
            }
        }
    };
    auto emit = [&](int c, int64_t sl, const Row &cv) __attribute__((always_inline)) {
        if (c < 0) return; // (wave-uniform)
        uint8_t *out = R8 + sl * dp;
        float ss = 0.f, s8 = 0.f;
#pragma unroll
        for (int t = 0; t < NV; ++t) {
            const int e0 = 4 * lane + 256 * t;
            if (FULL || e0 < dp) {
                float v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const float r = FULL || e0 + u < d ? qv[t][u] - cv[t][u] : 0.f;
                    ss = fmaf(r, r, ss);
                    v[u] = fminf(fmaxf(r * scale, -448.f), 448.f);
                }
                int wq = 0;
                wq = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], wq, false);
                wq = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], wq, true);
                *(int *)(out + e0) = wq;
                if (Rq) {
                    const float a0 = __builtin_amdgcn_cvt_f32_fp8(wq, 0), a1 = __builtin_amdgcn_cvt_f32_fp8(wq, 1);
                    const float a2 = __builtin_amdgcn_cvt_f32_fp8(wq, 2), a3 = __builtin_amdgcn_cvt_f32_fp8(wq, 3);
                    s8 = fmaf(a0, a0, fmaf(a1, a1, fmaf(a2, a2, fmaf(a3, a3, s8))));
                }
            }
        }
        ss = wave_sum_dpp(ss);
        if (lane == 0) Rn[sl] = ss;
        if (Rq) {
            s8 = wave_sum_dpp(s8);
            if (lane == 0) Rq[sl] = sqrtf(s8) * (1.f + 1.f / 1024.f);
        }
    };
    for (int pb = p_lo; pb < p_lo + npw; pb += 64) {
        const int nj = min(64, p_lo + npw - pb);
        const int myc = lane < nj ? probes[q * np + pb + lane] : -1;
        const int64_t mysl = lane < nj ? (int64_t)pslot[q * np + pb + lane] : 0;
        // the probes this wave converts, in order: j walks the set bits of rem (wave-uniform), the
        // next one's centroid row loaded while this one's residual is made (two register sets
        // rotated by name)
        uint64_t rem = __ballot(myc >= 0 && (!xsplit || (myc & 7) == xcd));
        if (!rem) continue;
        int j = __builtin_ctzll(rem);
        rem &= rem - 1;
        Row ca, cb;
        load_c(__shfl(myc, j, 64), ca);
        for (;;) {
            const int ja = rem ? __builtin_ctzll(rem) : j;
            if (rem) load_c(__shfl(myc, ja, 64), cb);
            emit(__shfl(myc, j, 64), __shfl(mysl, j, 64), ca);
            if (!rem) break;
            rem &= rem - 1;
            j = ja;
            const int jb = rem ? __builtin_ctzll(rem) : j;
            if (rem) load_c(__shfl(myc, jb, 64), ca);
            emit(__shfl(myc, j, 64), __shfl(mysl, j, 64), cb);
            if (!rem) break;
            rem &= rem - 1;
            j = jb;
        }
    }
}

// ---------------------------------------------------------------- segment bounds per query
// The list scan in fp8 keeps, per (query, probe), keys below a per-query threshold; a query's
// threshold is the key just above T, an upper bound of the approximation of at least k2 rows of
// its first probe (so no row above T is among the k2 candidates the rerank takes), and the scan
// skips the 32-row groups in which no query of a query group can find a row below its threshold.
// Both come from the list's segments (its parts, or the whole list): for every row x of a segment
// and query residual q8, |q8.x8 - q8.centre| <= |q8| E with E = R + 3 X 2^-11 (seg_meta_kernel;
// the 3 X 2^-11 covers the fp32 rounding of both products' sums), so the scan's approximation
// qv + xr + m2 q8.x8 (m2 < 0) of any row of the segment lies in
//   [qv + mxr + m2 (q8.centre + |q8| E), qv + Mxr + m2 (q8.centre - |q8| E)],
// each end moved outward by 2^-12 of its terms' magnitude for the fp32 rounding of both sides.
// One workgroup of four waves (one per 32-query tile) per work item (a list and up to 128 of its
// queries); the tile's query fragments are held in registers across the list's centre tiles (NC =
// dp / 64 chunks at most), lane (col, half) holding query col and segments 8 (r / 4) + 4 half + r % 4
// of each centre tile.
// MODE 0 (bounds): for the item's pairs of probe rank 0, T = the least segment upper bound u such
// that the segments whose upper bound is at most u hold at least k2 rows, into qthr[query]
// (queries without such a T keep the value the caller filled, kKey32None: no bound).
// MODE 1 (groups): bit G & 31 of act[(item 4 + qt) kPruneWords + (G >> 5) - (g0 >> 5)] is set when
// some query of tile qt may need group G (a segment overlapping G with its lower bound's key below
// the query's threshold), and act[any_at + item 4 + qt] says whether any group is needed. Tiles
// without a query write zeros; padding columns repeat the item's last query. Lists spanning more
// than kPruneWords words of groups are not pruned (their "any" word is set).
constexpr int kPruneWords = 64;
constexpr int kSegMax = 64; // segments per list at most (the parts of a list: kSubMax)
template <int NC, int MODE>
__global__ __launch_bounds__(256) void seg_bound_kernel(const uint8_t *__restrict__ C8,
                                                        const float4 *__restrict__ smeta,
                                                        const int64_t *__restrict__ soff,
                                                        const int32_t *__restrict__ sbase,
                                                        const int64_t *__restrict__ loff, int dp, float scale,
                                                        const uint8_t *__restrict__ R8, const float *__restrict__ Rn,
                                                        const float *__restrict__ Rq, const int4 *__restrict__ items,
                                                        const int32_t *__restrict__ bins,
                                                        const int32_t *__restrict__ pair_q,
                                                        const int32_t *__restrict__ pair_p,
                                                        uint32_t *__restrict__ qthr, uint32_t *__restrict__ act,
                                                        int64_t any_at, int k2) {
    __shared__ float s_ub[4][32][kSegMax]; // MODE 0: per wave, query column and segment: upper bound
    __shared__ int s_rows[kSegMax], s_ga[kSegMax], s_gb[kSegMax]; // per segment: rows, first and last group
    const int item = blockIdx.x;
    if (item >= bins[3 * kXcds]) return; // (the item total)
    const int4 it = items[item];
    const int c = it.x, start = it.y, cnt = it.z;
    const int sb0 = sbase[c], sb1 = min(sbase[c + 1], sb0 + kSegMax); // the list's segments (<= kSegMax)
    if ((int)threadIdx.x < sb1 - sb0) { // the segments' extents, read once for the workgroup
        const int64_t a = soff[sb0 + threadIdx.x], b = soff[sb0 + threadIdx.x + 1];
        s_rows[threadIdx.x] = (int)(b - a);
        s_ga[threadIdx.x] = (int)(a >> 5);
        s_gb[threadIdx.x] = (int)((b - 1) >> 5);
    }
    __syncthreads(); // (before any wave leaves)
    const int lane = threadIdx.x & 63, half = lane >> 5, col = lane & 31;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int q0 = 32 * w;
    const int r0 = (int)loff[c], r1 = (int)loff[c + 1];
    const int g0 = r0 >> 5, g1 = (r1 - 1) >> 5, cg0 = g0 >> 5, nwd = (g1 >> 5) - cg0 + 1;
    uint32_t *out = act + ((int64_t)item * 4 + w) * kPruneWords;
    uint32_t *any = act + any_at + (int64_t)item * 4 + w;
    if (MODE == 1 && nwd > kPruneWords) { // (the scan scans such lists whole)
        if (lane == 0) *any = 1u;
        return;
    }
    if (q0 >= cnt) {
        if (MODE == 1) {
            for (int wi = lane; wi < nwd; wi += 64) out[wi] = 0u;
            if (lane == 0) *any = 0u;
        }
        return;
    }
    const int64_t slot = start + min(q0 + col, cnt - 1); // the lane's query (pair slot)
    const bool first = pair_p[slot] == 0;               // (MODE 0: the pairs to bound)
    if (MODE == 0 && !__any(first)) return;
    const int nch = dp / kIvfKC;
    const float qv = Rn[slot], qn = Rq[slot];
    const uint32_t th = MODE == 1 ? qthr[pair_q[slot]] : 0u;
    const float m2 = -2.f / (scale * scale);
    V8 b[NC];
    const uint8_t *qrow = R8 + slot * dp + 32 * half;
#pragma unroll
    for (int ch = 0; ch < NC; ++ch)
        if (ch < nch) {
            b[ch].u[0] = *(const uint4 *)(qrow + ch * kIvfKC);
            b[ch].u[1] = *(const uint4 *)(qrow + ch * kIvfKC + 16);
        }
    uint64_t smask = 0;                            // MODE 1, bit s - sb0: segment s needed
    const uint8_t *lane_c8 = C8 + col * 64 + 32 * half;
    for (int ct = sb0 >> 5; ct <= (sb1 - 1) >> 5; ++ct) {
        const uint8_t *cb = lane_c8 + (((int64_t)ct * nch) << 11);
        V8 a[NC];
#pragma unroll
        for (int ch = 0; ch < NC; ++ch)
            if (ch < nch) {
                const uint4 *src = (const uint4 *)(cb + ((int64_t)ch << 11));
                a[ch].u[0] = src[0];
                a[ch].u[1] = src[1];
            }
        f32x16 acc = {};
#pragma unroll
        for (int ch = 0; ch < NC; ++ch)
            if (ch < nch)
                acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a[ch].v, b[ch].v, acc, 0, 0, 0, 127, 0, 127);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int sg = 32 * ct + 8 * (r >> 2) + 4 * half + (r & 3);
            const bool in = sg >= sb0 && sg < sb1;
            const float4 sv = smeta[min(max(sg, sb0), sb1 - 1)];
            const float e = qn * (sv.x + 3.f * sv.y * (1.f / 2048.f));
            if (MODE == 0) {
                const float U = acc[r] - e; // (<= q8.x8 of every row of the segment)
                float ub = fmaf(m2, U, qv + sv.z);
                ub += (qv + sv.z - m2 * fabsf(U)) * (1.f / 4096.f);
                if (in) s_ub[w][col][sg - sb0] = ub;
            } else {
                const float U = acc[r] + e; // (>= q8.x8 of every row of the segment)
                float lb = fmaf(m2, U, qv + sv.w);
                lb -= (qv + sv.w - m2 * fabsf(U)) * (1.f / 4096.f);
                const bool need = in && (f32_key(fmaxf(0.f, lb)) & ~15u) < th;
                const uint64_t bal = __ballot(need);
                const int s0 = 32 * ct + 8 * (r >> 2) + (r & 3) - sb0; // half 0's segment, relative
                if ((uint32_t)bal && s0 >= 0 && s0 < 64) smask |= 1ull << s0;
                if ((uint32_t)(bal >> 32) && s0 + 4 >= 0 && s0 + 4 < 64) smask |= 1ull << (s0 + 4);
            }
        }
    }
    if (MODE == 0) {
        // lane (col, half) tries the upper bounds of segments half, half + 2, ... of query col as T;
        // the two halves' least valid T is taken
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int ns = sb1 - sb0;
        // a segment of at least k2 rows is a valid T by itself: the least such bound first (one
        // pass), then only the smaller bounds need the count over all segments
        float best = INFINITY;
        for (int t = half; t < ns; t += 2)
            if (s_rows[t] >= k2) best = fminf(best, s_ub[w][col][t]);
        best = fminf(best, __shfl_xor(best, 32, 64));
        for (int t = half; t < ns; t += 2) {
            const float u = s_ub[w][col][t];
            if (u >= best) continue;
            int64_t rows = 0;
            for (int s2 = 0; s2 < ns; ++s2)
                if (s_ub[w][col][s2] <= u) rows += s_rows[s2];
            if (rows >= k2) best = fminf(best, u);
        }
        best = fminf(best, __shfl_xor(best, 32, 64));
        if (half == 0 && first && q0 + col < cnt && best < INFINITY)
            qthr[pair_q[slot]] = (f32_key(fmaxf(0.f, best)) & ~15u) + 16u;
        return;
    }
    // needed segments -> needed groups (word wi: groups 32 (cg0 + wi) ..)
    for (int wi = lane; wi < nwd; wi += 64) {
        const int wg0 = 32 * (cg0 + wi);
        uint32_t m32 = 0;
        for (uint64_t m = smask; m; m &= m - 1) {
            const int s = __builtin_ctzll(m);
            const int ga = s_ga[s] - wg0, gb = s_gb[s] - wg0;
            if (gb < 0 || ga > 31) continue;
            const int lo = max(ga, 0), hi = min(gb, 31);
            m32 |= (hi == 31 ? 0xFFFFFFFFu : ((2u << hi) - 1u)) & ~((1u << lo) - 1u);
        }
        out[wi] = m32;
    }
    if (lane == 0) *any = smask != 0 ? 1u : 0u;
}

// The list scan in fp8 e4m3 on residuals to the list's centroid c: |q - x|^2 =
// |q - c|^2 + |x - c|^2 - 2 (q - c).(x - c), so the quantised vectors are the short residuals.
// One workgroup of 8 waves per item of up to 128 queries. With more than 64 queries the waves
// form two query groups (64 queries each) of four row groups, so each fetched row serves 128
// queries (the second group's fetch of the same fragment comes from L2); with at most 64, all
// eight waves are row groups of one query group. The workgroup copies the item's query
// residuals (q - c) * scale (fp8, made for every pair by query_residual_kernel, with
// |q - c|^2 in fp32) into LDS once; the K loop then needs no staging: per 64-dim chunk one
// v_mfma_scale_f32_32x32x64_f8f6f4 (unit scales) per (row tile, query tile), lane l holding
// bytes 32 (l/32) .. +31 of row / query l % 32 of the chunk; the products are scaled back by
// 1 / scale^2. The list's rows are taken as whole 32-row groups of the fp8_row_chunk layout (a
// list's first and last groups may hold rows of its neighbours, which are masked), so a wave's
// tile and chunk name one 2 KB block at a wave-uniform address; fragments are loaded two chunks
// ahead, across row blocks. Keys: fp8_rowblock_keys (one threshold per query shared by its
// lists); lists and the final merge are the bf16 scan's (8 lists per query in two groups, 16 in
// one). PD: how many steps ahead the fragments are loaded (2 or 3). (Staging each step's rows once
// in LDS for both query groups, which halves the L2 reads, measured slower: a barrier per step.)
template <int KL, int PD>
__global__ __launch_bounds__(512, 1) void ivfl2_scan_fp8_kernel(
    const uint8_t *__restrict__ L8, const float *__restrict__ rn, const int64_t *__restrict__ loff, int dp,
    float scale, const uint8_t *__restrict__ R8, const float *__restrict__ Rn,
    const int4 *__restrict__ items, int32_t *__restrict__ bins, int nranges, const int32_t *__restrict__ pair_q,
    const int32_t *__restrict__ pair_p, int np, const int32_t *__restrict__ qoff, const uint32_t *__restrict__ qthr,
    uint64_t *__restrict__ cand, int nmax, const uint32_t *__restrict__ act, int64_t any_at,
    unsigned long long *__restrict__ work) {
    // work: per workgroup, the MFMA instructions its waves issued and the fp8 row bytes they fetched
    constexpr int IQ = kIvf8IQ;
    // IQ x (dp + 16) bytes of query residuals, then the list's row norms (nmax floats)
    extern __shared__ __align__(16) unsigned char dsm[];
    float *s_norm = (float *)(dsm + IQ * (dp + 16)); // the list's groups' row norms (<= nmax rows)
    __shared__ int32_t s_q[IQ];
    __shared__ uint32_t s_thr[IQ]; // per query: the smallest threshold any of its lists posted
    __shared__ float s_qr[IQ];
    __shared__ int s_slot[2];
    const int QS = dp + 16; // LDS row stride of the residual block (16 B of pad spreads the banks)
    const int tid = threadIdx.x, lane = tid & 63, half = lane >> 5, col = lane & 31;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6); // (wave-uniform: tile addresses in SGPRs)
    const float m2 = -2.f / (scale * scale);
    const int xb = nranges == 1 ? 0 : blockIdx.x % kXcds;
    // Work items are taken one ahead: thread 0 takes the next item when an item starts and posts
    // it in the other s_slot entry at the item's end, so the queue's atomic overlaps the item.
    // (one range: the queue's item count and first item are fixed for the launch; the atomic's
    // result is first needed at the item's end)
    const int cnt0 = bins[kXcds], first0 = bins[0];
    unsigned long long wk_mfma = 0, wk_bytes = 0; // (work[2 blockIdx.x ..]: this workgroup's work)
    if (tid == 0) s_slot[0] = take_item(bins, nranges, xb);
    for (int par = 0;; par ^= 1) {
        __syncthreads(); // (s_slot[par] posted; the previous item's LDS is no longer read)
        // (read into a scalar register: the item and every list bound and tile address derived
        // from it are wave-uniform, computed on the scalar unit)
        const int slot = __builtin_amdgcn_readfirstlane(s_slot[par]);
        if (slot < 0) break;
        int next = -1;
        if (tid == 0) next = nranges == 1 ? atomicAdd(&bins[2 * kXcds], 1) : take_item(bins, nranges, xb);
        const int4 it = items[slot];
        const int c = it.x, start = it.y, cnt = it.z;
        // rows and groups in 32 bits (the database has fewer than 2^31 rows)
        const int r0 = (int)loff[c], r1 = (int)loff[c + 1];
        if (act) {
            // an item no query of which needs any group (prune_groups_kernel) keeps no candidate
            const uint4 an = *(const uint4 *)(act + any_at + (int64_t)slot * 4);
            if ((an.x | an.y | an.z | an.w) == 0u) {
                for (int e = tid; e < cnt * kIvf8Cap; e += 512) {
                    const int t = e / kIvf8Cap;
                    cand[cand_slot(qoff, pair_q[start + t], np, pair_p[start + t]) * kIvf8Cap + e % kIvf8Cap] =
                        kKey64None;
                }
                if (tid == 0) s_slot[par ^ 1] = nranges == 1 ? (next < cnt0 ? first0 + next : -1) : next;
                continue;
            }
        }
        const bool wide = cnt > 64;
        const int qg = wide ? w >> 2 : 0, wr = wide ? w & 3 : w; // query group, row group
        const int nrg = wide ? 4 : 8, nlist = 2 * nrg, tstride = wide ? IQ : 64;
        const int q0 = 64 * qg;                 // the wave's first query
        const bool two = cnt - q0 > 32;         // (else its second query tile is padding)
        const int nrows = wide ? IQ : 64;       // query rows in LDS
        if (tid < IQ) {
            s_q[tid] = pair_q[start + min(tid, cnt - 1)];
            s_thr[tid] = qthr ? qthr[pair_q[start + min(tid, cnt - 1)]] : kKey32None;
        }
        { // the item's query residuals (pair slots start ..) into LDS, four loads in flight per
          // thread; padding rows repeat the last
            const int upr = dp / 16, tot = nrows * upr;
            for (int u0 = tid; u0 < tot; u0 += 4 * 512) {
                uint4 v[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int u = min(u0 + 512 * k, tot - 1), t = u / upr, e = u - t * upr;
                    v[k] = *(const uint4 *)(R8 + (int64_t)(start + min(t, cnt - 1)) * dp + 16 * e);
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int u = u0 + 512 * k, t = u / upr, e = u - t * upr;
                    if (u < tot) *(uint4 *)(dsm + t * QS + 16 * e) = v[k];
                }
            }
            if (tid < nrows) s_qr[tid] = Rn[start + min(tid, cnt - 1)];
        }
        // the list's 32-row groups g0 .. g1; a row block is 2 nrg of them, two per wave
        const int g0 = r0 >> 5, g1 = (r1 - 1) >> 5;
        const int ngr = g1 - g0 + 1;
        const int nch = dp / kIvfKC;
        const int GB = 2 * nrg; // groups per row block
        // the list is scanned in segments of whole row blocks whose row norms fit in LDS (nmax is
        // a multiple of 512 rows, 16 groups: every segment but the last is whole row blocks); a
        // list of at most nmax rows is one segment
        const int segg = nmax / 32;
        const uint8_t *lane_l8 = L8 + col * 64 + 32 * half;
        uint32_t Lk[2][KL], Li[2][KL];
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < KL; ++e) {
                Lk[j][e] = kKey32None;
                Li[j][e] = 0u;
            }
        bool fresh = true; // (the item's lists are empty until its first row block's keys)
        for (int sg = 0; sg < ngr; sg += segg) {
            const int sng = min(segg, ngr - sg);
            const int sg0 = g0 + sg, sg1 = sg0 + sng - 1; // the segment's groups
            if (sg > 0) __syncthreads(); // (every wave has read the previous segment's norms)
            for (int e0 = tid; e0 < 32 * sng; e0 += 4 * 512) {
                float v[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) v[k] = rn[32 * sg0 + min(e0 + 512 * k, 32 * sng - 1)];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (e0 + 512 * k < 32 * sng) s_norm[e0 + 512 * k] = v[k];
            }
            __syncthreads(); // (and, for the first segment, the item's query block)
            const float qr0 = s_qr[q0 + col], qr1 = s_qr[q0 + 32 + col];
            const int nrb = (sng + GB - 1) / GB; // (<= 32: segg <= 256 groups, GB >= 8)
            // the wave's row blocks to scan: all, or with group pruning (act) those with a group that
            // some query of the wave's query group may need
            uint32_t rbmask = nrb >= 32 ? 0xFFFFFFFFu : ((1u << nrb) - 1u);
            uint32_t uniq = qg == 0 ? rbmask : 0u; // (work count: blocks whose rows this wave reads first)
            if (act && (g1 >> 5) - (g0 >> 5) < kPruneWords) {
                // groups that some query of the wave's two query tiles may need (seg_bound_kernel):
                // lane l tests row block l
                const int cgl = g0 >> 5;
                auto mask_of = [&](int g) { // query group g's row blocks
                    const uint32_t *aq = act + ((int64_t)slot * 4 + 2 * g) * kPruneWords;
                    bool on = false;
                    if (lane < nrb)
#pragma unroll
                        for (int i = 0; i < 2; ++i) {
                            const int G = sg0 + GB * lane + 2 * wr + i;
                            if (G <= sg1) {
                                const int wi = (G >> 5) - cgl;
                                on = on || (((aq[wi] | aq[kPruneWords + wi]) >> (G & 31)) & 1u);
                            }
                        }
                    return (uint32_t)__ballot(on);
                };
                rbmask = mask_of(qg);
                // (the rows of blocks that query group 0 also scans come from L2: not counted twice)
                if (work && qg == 1) uniq = rbmask & ~mask_of(0);
            }
            if (act && qg == 0) uniq = rbmask;
            { // the wave's products and the row bytes the item reads (summed per workgroup at the end)
                wk_mfma += (unsigned long long)__popc(rbmask) * (unsigned long long)(nch * (two ? 4 : 2));
                wk_bytes += (unsigned long long)__popc(uniq) * 64ull * (unsigned long long)dp;
            }
            if (rbmask == 0) continue; // (wave-uniform; no barrier follows inside the segment)
            // the active row block after b, or -1
            auto next_rb = [&](int b) { return b >= 31 ? -1 : __ffs(rbmask & (0xFFFFFFFEu << b)) - 1; };
            // row fragments of step g (the g / nch-th active row block, chunk g % nch), loaded PD
            // steps ahead into a ring of PD + 1 register sets that rotate by name (PD + 1 steps per
            // loop trip): a copy between sets would wait for the loads in flight. Groups past the
            // segment's last are clamped to it (they lie past g1: their keys are skipped).
            V8 fa[2], fb[2], fc[2];
            const int nsteps = __popc(rbmask) * nch;
            // the loads walk the steps in order, each trip's last ones clamped to the final step:
            // a cursor (row block lrb, chunk lch) replaces a division per load
            int lrb = __ffs(rbmask) - 1, lch = 0;
            const int last_rb = 31 - __clz(rbmask), last_ch = nch - 1;
            auto load_a = [&](V8 (&a)[2]) {
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const int64_t grp = min(sg0 + GB * lrb + 2 * wr + i, sg1);
                    const uint4 *src = (const uint4 *)(lane_l8 + ((grp * nch + lch) << 11));
                    a[i].u[0] = src[0];
                    a[i].u[1] = src[1];
                }
                if (lrb < last_rb || lch < last_ch) {
                    if (++lch == nch) {
                        lch = 0;
                        lrb = next_rb(lrb);
                    }
                }
            };
            f32x16 acc[2][2] = {}; // (defined before the loop: no undefined values live across it)
            int ch = 0, rb = __ffs(rbmask) - 1;
            // one chunk: its products, and after a row block's last chunk its keys
            auto chunk = [&](const V8 (&cur)[2]) {
                if (ch == 0)
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int j = 0; j < 2; ++j) acc[i][j] = f32x16{};
                V8 b[2];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const uint4 *sb = (const uint4 *)(dsm + (q0 + 32 * j + col) * QS + ch * kIvfKC + 32 * half);
                    b[j].u[0] = sb[0];
                    b[j].u[1] = sb[1];
                }
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        if (j == 0 || two)
                            acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(cur[i].v, b[j].v, acc[i][j], 0,
                                                                                   0, 0, 127, 0, 127);
                if (++ch == nch) {
                    const int grp = sg0 + GB * rb + 2 * wr; // the wave's first group of the block
                    // (a row block past the segment, from a trip's last steps, is skipped whole)
                    fp8_rowblock_keys<KL>(acc, m2, s_norm + 32 * (grp - sg0), 32 * grp, r0, min(r1, 32 * (sg1 + 1)),
                                          fresh, qr0, qr1, Lk, Li, s_thr, q0 + col, half, two);
                    fresh = false;
                    ch = 0;
                    const int nx = next_rb(rb);
                    rb = nx < 0 ? nrb : nx; // (past the last active block: nrb, a block past the segment)
                }
            };
            // one step: the load PD steps ahead, then the chunk. Whole trips: up to PD steps past
            // the end run on clamped rows of the last chunk and end in a row block past the
            // segment, whose keys are skipped. (With three steps ahead in four sets, the compiler's
            // waits, which across the trip's branches keep one step fewer in flight than the ring
            // allows, still leave two.)
            auto step = [&](const V8 (&cur)[2], V8 (&ahead)[2]) {
                load_a(ahead); // step min(g + PD, nsteps - 1) (unconditional, so the waits stay exact)
                chunk(cur);
            };
            load_a(fa);
            load_a(fb);
            if constexpr (PD == 2) {
                for (int g = 0; g < nsteps; g += 3) {
                    step(fa, fc);
                    step(fb, fa);
                    step(fc, fb);
                }
            } else {
                V8 fd[2];
                load_a(fc);
                for (int g = 0; g < nsteps; g += 4) {
                    step(fa, fd);
                    step(fb, fa);
                    step(fc, fb);
                    step(fd, fc);
                }
            }
        }
        // the lists go to LDS (the residual block is no longer read), then each query's nlist x KL
        // entries are sorted across one wave and its kIvf8Cap = 64 best go to its (query, probe)
        // slot
        __syncthreads();
        uint64_t *sl = (uint64_t *)dsm; // [list][query (tstride)][KL]
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < KL; ++e)
                sl[((wr * 2 + half) * tstride + q0 + 32 * j + col) * KL + e] =
                    Lk[j][e] == kKey32None ? kKey64None : (((uint64_t)Lk[j][e] << 32) | Li[j][e]);
        __syncthreads();
        // (each lane list is sorted: odd lists are read backwards, so the runs alternate)
        auto entry = [&](int m, int t) {
            const int l = m / KL, e = m % KL;
            return sl[(l * tstride + t) * KL + ((l & 1) ? KL - 1 - e : e)];
        };
        // (a query whose lists are all empty, as under a bound most are, skips the network: the
        // sorts are chains of dependent shuffles, as long as a good part of an item's products)
        for (int t = w; t < cnt; t += 8) {
            uint64_t *out = cand + cand_slot(qoff, s_q[t], np, pair_p[start + t]) * kIvf8Cap;
            if (nlist * KL <= 64) { // two query groups (8 lists), KL = 8: one entry per lane
                uint64_t v[1] = {entry(lane, t)};
                if (__any(v[0] != kKey64None)) wave_sort_u64<1, 2 * KL>(v, lane);
                out[lane] = v[0];
            } else {
                constexpr int E = KL / 4; // 16 lists x KL entries, E per lane
                uint64_t v[E];
                bool any = false;
#pragma unroll
                for (int s2 = 0; s2 < E; ++s2) {
                    const int m = lane + 64 * s2;
                    v[s2] = m < nlist * KL ? entry(m, t) : kKey64None;
                    any = any || v[s2] != kKey64None;
                }
                if (__any(any)) wave_sort_u64<E, 2 * KL>(v, lane);
                out[lane] = v[0]; // the 64 best
            }
        }
        if (tid == 0) s_slot[par ^ 1] = nranges == 1 ? (next < cnt0 ? first0 + next : -1) : next;
        // (the next item's LDS writes follow the loop's first barrier, which every wave reaches
        // only after this item's sorts)
    }
    if (work) { // (every wave left the loop after the same barrier: the LDS below is free)
        unsigned long long *s_wk = (unsigned long long *)dsm;
        if (lane == 0) {
            s_wk[2 * w] = wk_mfma;
            s_wk[2 * w + 1] = wk_bytes;
        }
        __syncthreads();
        if (tid < 2) {
            unsigned long long t = 0;
            for (int v = 0; v < 8; ++v) t += s_wk[2 * v + tid];
            work[2 * blockIdx.x + tid] = t;
        }
    }
}

// The list scan. One workgroup (4 waves) per work item: up to 64 queries of one cluster against
// all rows of that cluster. Rows go in blocks of 256 (64 per wave, two 32-row MFMA tiles); the
// queries are two 32-query tiles, so a wave holds 2 x 2 accumulators of 32 x 32. The K loop runs
// in chunks of 64 dims: the query chunk is staged in LDS (two buffers, the next chunk is loaded
// into registers during the current chunk's products), the wave's row fragments are read straight
// from HBM/L2 into registers, issued for the next chunk right after the current chunk's products
// (for the next row block's first chunk: after this row block's keys).
// dp is a multiple of the chunk (64). After a
// row block, lane (c, h) of wave w turns its 32 accumulators of query c (+ 32 j) into keys
// |q|^2 + |x|^2 - 2 q.x and keeps the KL smallest (key, row position) per query; at the end the
// 8 lists (4 waves x 2 halves) of each query are merged into its KP best for this probe.
template <int PREC, int KL>
__global__ __launch_bounds__(256, 2) void ivfl2_scan_kernel(
    const __bf16 *__restrict__ Lh, const __bf16 *__restrict__ Ll, const float *__restrict__ lxn,
    const int64_t *__restrict__ loff, int dp, const __bf16 *__restrict__ Qh, const __bf16 *__restrict__ Ql,
    const float *__restrict__ qn, const int4 *__restrict__ items, int32_t *__restrict__ bins, int nranges,
    const int32_t *__restrict__ pair_q, const int32_t *__restrict__ pair_p, int np, const int32_t *__restrict__ qoff,
    const uint32_t *__restrict__ qthr, uint64_t *__restrict__ cand, int kp) {
    constexpr int QS = kIvfQS, IQ = kIvfIQ;
    constexpr int NPL = PREC == 3 ? 2 : 1;      // staged planes: hi (and lo)
    constexpr int PLANE = IQ * QS;              // bf16 elements of one staged plane
    constexpr int BUF = PLANE * NPL;            // one query-chunk buffer
    constexpr int SCAN_BYTES = 2 * BUF * 2;
    constexpr int LIST_BYTES = 8 * IQ * KL * 8;
    __shared__ __align__(16) unsigned char smem[SCAN_BYTES > LIST_BYTES ? SCAN_BYTES : LIST_BYTES];
    __shared__ int32_t s_q[IQ];
    __shared__ uint32_t s_thr[8][IQ]; // per (wave, half) list: its query's prune threshold
    __bf16 *sq = (__bf16 *)smem;

    __shared__ int s_slot;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, half = lane >> 5, col = lane & 31;
    // workgroups take the items of their range (per XCD: id % kXcds; else the one range) in order,
    // then those left in the other ranges
    const int xb = nranges == 1 ? 0 : blockIdx.x % kXcds;
    for (;;) {
    if (tid == 0) s_slot = take_item(bins, nranges, xb);
    __syncthreads();
    const int slot = __builtin_amdgcn_readfirstlane(s_slot); // (wave-uniform item: scalar addressing)
    if (slot < 0) break;
    const int4 it = items[slot];
    const int c = it.x, start = it.y, cnt = it.z;
    const bool two = cnt > 32; // (else the second query tile is padding)
    if (tid < IQ) s_q[tid] = pair_q[start + min(tid, cnt - 1)];
    for (int e = tid; e < 8 * IQ; e += 256)
        s_thr[e / IQ][e % IQ] = qthr ? qthr[pair_q[start + min(e % IQ, cnt - 1)]] : kKey32None;
    __syncthreads();
    const int64_t r0 = loff[c], r1 = loff[c + 1];
    const int nch = dp / kIvfKC;
    const int nsteps = (int)((r1 - r0 + 255) / 256) * nch;
    // staging role: 16-byte units tid and tid + 256 -> (query su, su + 32; dims 8 sp .. 8 sp + 7)
    const int su = tid >> 3, sp = tid & 7;
    const int64_t sg0 = (int64_t)s_q[su] * dp + 8 * sp, sg1 = (int64_t)s_q[su + 32] * dp + 8 * sp;
    const float qn0 = qn[s_q[col]], qn1 = qn[s_q[col + 32]];
    const float dscale0 = 1.f, dscale1 = 1.f; // the products are the dot products themselves

    bf16x8 qreg[2][NPL];
    auto load_q = [&](int kc) {
#pragma unroll
        for (int pl = 0; pl < NPL; ++pl) {
            const __bf16 *src = pl == 0 ? Qh : Ql;
            qreg[0][pl] = *(const bf16x8 *)(src + sg0 + kc);
            qreg[1][pl] = *(const bf16x8 *)(src + sg1 + kc);
        }
    };
    auto store_q = [&](int buf) {
#pragma unroll
        for (int pl = 0; pl < NPL; ++pl) {
            *(bf16x8 *)(sq + buf * BUF + pl * PLANE + su * QS + 8 * sp) = qreg[0][pl];
            *(bf16x8 *)(sq + buf * BUF + pl * PLANE + (su + 32) * QS + 8 * sp) = qreg[1][pl];
        }
    };
    // row fragments of (row block rb, chunk ch): rows r0 + 256 rb + 64 w + 32 i + col (clamped
    // into the list), dims 64 ch + 16 s + 8 half
    bf16x8 acur[2][4][NPL];
    auto load_a = [&](int rb, int ch, bf16x8 (&a)[2][4][NPL]) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int64_t row = min(r0 + 256 * rb + 64 * w + 32 * i + col, r1 - 1);
#pragma unroll
            for (int pl = 0; pl < NPL; ++pl) {
                const __bf16 *src = (pl == 0 ? Lh : Ll) + row * dp + ch * kIvfKC + 8 * half;
#pragma unroll
                for (int s = 0; s < 4; ++s) a[i][s][pl] = *(const bf16x8 *)(src + 16 * s);
            }
        }
    };

    uint32_t Lk[2][KL], Li[2][KL];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int e = 0; e < KL; ++e) {
            Lk[j][e] = kKey32None;
            Li[j][e] = 0u;
        }
    load_q(0);
    store_q(0);
    load_a(0, 0, acur);
    __syncthreads();
    const int nrb = (int)((r1 - r0 + 255) / 256);
    int g = 0;
    for (int rb = 0; rb < nrb; ++rb) {
        f32x16 acc[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = f32x16{};
        for (int ch = 0; ch < nch; ++ch, ++g) {
            const int buf = g & 1;
            const bool more = g + 1 < nsteps;
            const int ch1 = ch + 1 == nch ? 0 : ch + 1;
            if (more) load_q(ch1 * kIvfKC);
            const __bf16 *qb = sq + buf * BUF;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                bf16x8 b[2][NPL];
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int pl = 0; pl < NPL; ++pl)
                        b[j][pl] = *(const bf16x8 *)(qb + pl * PLANE + (32 * j + col) * QS + 16 * s + 8 * half);
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        if (j == 1 && !two) continue;
                        acc[i][j] =
                            __builtin_amdgcn_mfma_f32_32x32x16_bf16(acur[i][s][0], b[j][0], acc[i][j], 0, 0, 0);
                        if (PREC == 3) {
                            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(acur[i][s][0], b[j][NPL - 1],
                                                                                acc[i][j], 0, 0, 0);
                            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(acur[i][s][NPL - 1], b[j][0],
                                                                                acc[i][j], 0, 0, 0);
                        }
                    }
            }
            if (more) store_q(buf ^ 1);
            if (ch + 1 < nch) load_a(rb, ch + 1, acur); // the next row block's first chunk: after the keys
            __syncthreads();
        }
        rowblock_keys<KL>(acc, dscale0, dscale1, lxn, r0 + 256 * rb + 64 * w, r1, rb == 0, qn0, qn1, Lk, Li,
                          &s_thr[0][0], IQ, 8, w * 2 + half, col, half, two);
        if (rb + 1 < nrb) load_a(rb + 1, 0, acur);
    }
    merge_item_lists<KL, 8>(smem, Lk, Li, s_q, pair_p + start, cnt, np, qoff, cand, IQ, 8, w * 2 + half, col, kp);
    __syncthreads(); // LDS (lists, query chunks, s_q, s_thr, s_slot) is reused by the next item
    }
}

__device__ __forceinline__ uint64_t wmin64(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}

// Rerank order: queries grouped by their first probed list (a counting sort), so that queries
// sharing candidate rows are reranked together. key = the list of the first valid probe slot
// (nl: none).

// Exclusive scan of hist[0 .. m) in place (one workgroup of 256).
__global__ __launch_bounds__(256) void order_scan_kernel(int32_t *__restrict__ hist, int m) {
    __shared__ int32_t part[256];
    const int tid = threadIdx.x, per = (m + 255) / 256;
    const int o0 = min(m, tid * per), o1 = min(m, o0 + per);
    int sum = 0;
    for (int o = o0; o < o1; ++o) sum += hist[o];
    part[tid] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int a = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += a;
        __syncthreads();
    }
    int run = part[tid] - sum;
    for (int o = o0; o < o1; ++o) {
        const int v = hist[o];
        hist[o] = run;
        run += v;
    }
}


// sum (q - x)^2 over d dims by 16 lanes (c = lane % 16; every lane of the 16 gets the sum): with
// VEC, lane c takes the float4 pieces 4 c, 4 c + 64, ... into four chains joined as (0+1)+(2+3);
// otherwise dims c, c + 16, ... in one chain; then a butterfly over the 16 lanes. x = nullptr
// gives 0 (the lanes still take part in the butterfly). Every fp32 distance the IVF-L2 search
// returns is this sum.
template <bool VEC>
__device__ __forceinline__ float row_sq_dist16(const float *__restrict__ qrow, const float *__restrict__ xr,
                                               int64_t d, int c) {
    float part = 0.f;
    if (xr) {
        if (VEC) {
            float p0 = 0.f, p1 = 0.f, p2 = 0.f, p3 = 0.f;
#pragma unroll 6
            for (int64_t i = 4 * c; i < d; i += 64) {
                const float4 a = *(const float4 *)(qrow + i), b = *(const float4 *)(xr + i);
                const float t0 = a.x - b.x, t1 = a.y - b.y, t2 = a.z - b.z, t3 = a.w - b.w;
                p0 = fmaf(t0, t0, p0);
                p1 = fmaf(t1, t1, p1);
                p2 = fmaf(t2, t2, p2);
                p3 = fmaf(t3, t3, p3);
            }
            part = (p0 + p1) + (p2 + p3);
        } else {
            for (int64_t i = c; i < d; i += 16) {
                const float t = qrow[i] - xr[i];
                part = fmaf(t, t, part);
            }
        }
    }
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) part += __shfl_xor(part, off, 16);
    return part;
}

// One wave per query: the K2 best candidates by the approximation are recomputed in fp32 as
// sum (q - x)^2 by 16 lanes each (four candidates per round; with VEC, 16-byte loads, so a wave
// reads four rows per step in 256-byte pieces), then the k smallest (distance, id) are written
// ascending. X holds the rows in cluster order (a candidate's position indexes it; lids maps a
// position to its id), so the rows that queries of one cluster share are near each other.
template <bool VEC>
__global__ __launch_bounds__(256) void ivfl2_rerank_kernel(const float *__restrict__ X, int64_t d,
                                                           const int32_t *__restrict__ lids, const float *__restrict__ Q,
                                                           int64_t qstride, const uint64_t *__restrict__ cand, int np,
                                                           int cap, const int32_t *__restrict__ qoff,
                                                           int64_t nq, int k, int32_t *__restrict__ out_idx,
                                                           float *__restrict__ out_dist, int32_t *__restrict__ out_cnt,
                                                           const int32_t *__restrict__ qorder, int64_t per_xcd, int k2,
                                                           int32_t *__restrict__ ohist, int nh) {
    const int lane = threadIdx.x & 63;
    // the query order's histogram (read before this launch) is left zero for the next search
    for (int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x; b < nh; b += (int64_t)gridDim.x * 256) ohist[b] = 0;
    // workgroup b runs on XCD b % 8: each XCD takes a contiguous run of qorder
    const int64_t pos = ((int64_t)(blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds) * 4 + (threadIdx.x >> 6);
    if (pos >= nq) return;
    const int64_t q = qorder[pos];
    const uint64_t *src = cand + cand_slot(qoff, q, np, 0) * cap;
    const int npair = qoff ? qoff[q + 1] - qoff[q] : np;
    // Each (query, probe) slot holds its candidates ascending, empty entries last (the scans merge
    // them so). The query's 64 best so far are held across the wave, lane r holding the r-th
    // (`mine`, ascending over lanes); a slot whose first entry is below the current k2-th best is
    // merged in: its first min(cap, 64) entries, reversed across the lanes, against `mine` -- the
    // lane-wise minimum is a bitonic sequence of the 64 smallest of both, which six compare-exchange
    // stages sort. Slots that the scan's bounds left empty cost one read.
    uint64_t mine = kKey64None;
    const int mc = min(cap, 64);
    auto merge = [&](uint64_t b) { // b: lane j holds entry j of a slot (ascending)
        const uint64_t rb = __shfl(b, 63 - lane, 64);
        uint64_t c = rb < mine ? rb : mine;
#pragma unroll
        for (int j = 32; j > 0; j >>= 1) {
            const uint64_t o = __shfl_xor(c, j, 64);
            const bool lower = (lane & j) == 0;
            c = lower ? (o < c ? o : c) : (o < c ? c : o);
        }
        mine = c;
    };
    for (int p0 = 0; p0 < npair; p0 += 64) {
        const uint64_t first = p0 + lane < npair ? src[(int64_t)(p0 + lane) * cap] : kKey64None;
        uint64_t bal = __ballot(first != kKey64None);
        // two slots per round: both loads in flight before the merges
        while (bal) {
            const int pa = __builtin_ctzll(bal);
            bal &= bal - 1;
            const int pb = bal ? __builtin_ctzll(bal) : -1;
            if (pb >= 0) bal &= bal - 1;
            const uint64_t ba = lane < mc ? src[(int64_t)(p0 + pa) * cap + lane] : kKey64None;
            const uint64_t bb = pb >= 0 && lane < mc ? src[(int64_t)(p0 + pb) * cap + lane] : kKey64None;
            if (__shfl(first, pa, 64) < __shfl(mine, k2 - 1, 64)) merge(ba);
            if (pb >= 0 && __shfl(first, pb, 64) < __shfl(mine, k2 - 1, 64)) merge(bb);
        }
    }
    if (lane >= k2) mine = kKey64None; // (the k2 best are reranked)
    const int32_t mypos = mine == kKey64None ? -1 : (int32_t)(uint32_t)(mine & 0xFFFFFFFFull);
    const int32_t myid = mypos < 0 ? -1 : lids[mypos];
    const float *qrow = Q + q * qstride;
    uint64_t key = kKey64None;
    const int g = lane >> 4, c = lane & 15;
#pragma unroll 1
    for (int round = 0; round < (k2 + 3) / 4; ++round) {
        const int32_t pos = __shfl(mypos, 4 * round + g, 64); // group g: candidate 4 round + g
        const float part = row_sq_dist16<VEC>(qrow, pos >= 0 ? X + (int64_t)pos * d : nullptr, d, c);
        const bool owner = lane >= 4 * round && lane < 4 * round + 4;
        const float v = __shfl(part, owner ? 16 * (lane - 4 * round) : 0, 64);
        if (owner && myid >= 0) key = ((uint64_t)f32_key(v) << 32) | (uint32_t)myid;
    }
    int cnt = 0;
    for (int r = 0; r < k; ++r) {
        const uint64_t m = wmin64(key);
        if (m == kKey64None) break;
        if (key == m) key = kKey64None;
        if (lane == 0) {
            out_idx[q * k + r] = (int32_t)(uint32_t)(m & 0xFFFFFFFFull);
            out_dist[q * k + r] = key_f32((uint32_t)(m >> 32));
        }
        ++cnt;
    }
    if (lane == 0) {
        for (int r = cnt; r < k; ++r) {
            out_idx[q * k + r] = -1;
            out_dist[q * k + r] = -1.f;
        }
        out_cnt[q] = cnt;
    }
}

__global__ void gather_rows_f32_kernel(const float *__restrict__ X, int64_t d, const int32_t *__restrict__ ids,
                                       int64_t n, float *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n * d) out[t] = X[(int64_t)ids[t / d] * d + t % d];
}

// ---------------------------------------------------------------- row order inside a list
// A list's rows are ordered by a secondary partition of the list, ids ascending inside a part:
// kSubIters Lloyd iterations over m = min(kSubMax, n_list / kSubRows) part centroids seeded with
// evenly spaced rows of the list (lists of fewer than 2 kSubRows rows, and rows of more than
// kSubDims dims, keep id order). Candidates are named by id and reranked in fp32, so the order
// changes no result beyond ties of the approximation; it keeps the rows of a 32-row group close
// together, which lets the group's fp8 centre and radius bound its distances (group pruning in the
// fp8 scan). All sums run in a fixed order: the order is a function of the rows and lists alone.
constexpr int kSubRows = 384;
constexpr int kSubMax = 64;
constexpr int kSubIters = 8;
constexpr int kSubDims = 2048;

// the list of row `row` (rows in list order): the last list with loff[list] <= row
__device__ __forceinline__ int list_of_row(const int64_t *__restrict__ loff, int nl, int64_t row) {
    int lo = 0, hi = nl - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (loff[mid] <= row) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Part centroid p of list l starts as the list's row at (2 j + 1) n_l / (2 m) (j = p - pbase[l]).
__global__ __launch_bounds__(256) void sub_init_kernel(const float *__restrict__ X, int d,
                                                       const int64_t *__restrict__ loff,
                                                       const int32_t *__restrict__ plist,
                                                       const int32_t *__restrict__ pbase, float *__restrict__ P) {
    const int part = blockIdx.x, l = plist[part];
    const int64_t j = part - pbase[l], m = pbase[l + 1] - pbase[l];
    const int64_t nlr = loff[l + 1] - loff[l];
    const float *x = X + (loff[l] + (2 * j + 1) * nlr / (2 * m)) * d;
    for (int e = threadIdx.x; e < d; e += 256) P[(int64_t)part * d + e] = x[e];
}

// One wave per row (list order): the nearest part centroid of its list (lower index on ties);
// part 0 for lists without parts.
__global__ __launch_bounds__(256) void sub_assign_kernel(const float *__restrict__ X, int64_t n, int d,
                                                         const int64_t *__restrict__ loff, int nl,
                                                         const int32_t *__restrict__ pbase,
                                                         const float *__restrict__ P, int32_t *__restrict__ sub) {
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const int l = list_of_row(loff, nl, row);
    const int b = pbase[l], m = pbase[l + 1] - b;
    int bj = 0;
    if (m > 1) {
        const float *x = X + row * d;
        float best = INFINITY;
        for (int j = 0; j < m; ++j) {
            const float *p = P + (int64_t)(b + j) * d;
            float s = 0.f;
            for (int e = lane; e < d; e += 64) {
                const float t = x[e] - p[e];
                s = fmaf(t, t, s);
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
            s = __shfl(s, 0, 64); // (lane 0's sum: every lane compares the same value)
            if (s < best) {
                best = s;
                bj = j;
            }
        }
    }
    if (lane == 0) sub[row] = bj;
}

// One workgroup per part: the mean of the list's rows assigned to it, rows taken in ascending
// order (a part that holds no row keeps its centroid). d <= kSubDims.
__global__ __launch_bounds__(256) void sub_update_kernel(const float *__restrict__ X, int d,
                                                         const int64_t *__restrict__ loff,
                                                         const int32_t *__restrict__ plist,
                                                         const int32_t *__restrict__ pbase,
                                                         const int32_t *__restrict__ sub, float *__restrict__ P) {
    __shared__ int32_t s_rows[256];
    __shared__ int s_wc[4];
    const int part = blockIdx.x, l = plist[part];
    const int j = part - pbase[l];
    const int64_t r0 = loff[l], r1 = loff[l + 1];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    constexpr int T = kSubDims / 256;
    float acc[T];
#pragma unroll
    for (int t = 0; t < T; ++t) acc[t] = 0.f;
    int64_t cnt = 0;
    for (int64_t c0 = r0; c0 < r1; c0 += 256) {
        const int64_t r = c0 + tid;
        const bool hit = r < r1 && sub[r] == j;
        const uint64_t bal = __ballot(hit);
        if (lane == 0) s_wc[w] = __popcll(bal);
        __syncthreads();
        int off = 0;
        for (int u = 0; u < w; ++u) off += s_wc[u];
        const int nn = s_wc[0] + s_wc[1] + s_wc[2] + s_wc[3];
        if (hit) s_rows[off + __popcll(bal & ((1ull << lane) - 1))] = (int32_t)(r - r0);
        __syncthreads();
        for (int i = 0; i < nn; ++i) {
            const float *x = X + (r0 + s_rows[i]) * d;
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int e = tid + 256 * t;
                if (e < d) acc[t] += x[e];
            }
        }
        cnt += nn;
        __syncthreads();
    }
    if (cnt > 0)
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int e = tid + 256 * t;
            if (e < d) P[(int64_t)part * d + e] = acc[t] / (float)cnt;
        }
}

// Reorders the rows of each list (ids: list-order position -> id, X in that order) by the
// secondary partition; rewrites ids (host and device) and regathers X from the original rows dX.
static void sub_partition_order(const float *dX, int64_t n, int64_t d, const std::vector<int64_t> &off,
                                std::vector<int32_t> &ids, int32_t *dids, const int64_t *dloff, float *Xo,
                                std::vector<int64_t> &soff, std::vector<int32_t> &sbase, hipStream_t st);

// ---------------------------------------------------------------- segment bounds of the fp8 rows
// One wave per segment sg (rows soff[sg] .. soff[sg + 1] - 1 of the fp8 rows L8, fp8_row_chunk
// layout, dp bytes per row): the centre, the segment's rows' fp8 values averaged and rounded to
// fp8 as the rows are, goes to centre row sg of C8 (the same layout); sm[sg] = {R, X, Mxr, mxr}
// with R >= max |x8 - centre| and X >= max |x8| over the segment's rows (their fp32 sums rounded
// up by 2^-10) and Mxr / mxr the largest / least rn over them. With these, for every row x of the
// segment and any query residual q8, |q8.x8 - q8.centre| <= |q8| R, which bounds the scan's
// approximation of the segment's rows from both sides (seg_bound_kernel); |centre| <= 1.0625 X
// (the mean's norm is at most X before its fp8 rounding).
template <int NS>
__global__ __launch_bounds__(256) void seg_meta_kernel(const uint8_t *__restrict__ L8, const float *__restrict__ rn,
                                                       const int64_t *__restrict__ soff, int64_t nseg, int dp,
                                                       uint8_t *__restrict__ C8, float4 *__restrict__ sm) {
    const int64_t sg = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (sg >= nseg) return;
    const int64_t s0 = soff[sg], s1 = soff[sg + 1];
    const int nch = dp / kIvfKC;
    auto at = [&](int64_t row, int b) { return fp8_row_chunk(row, b / kIvfKC, nch) + b % kIvfKC; };
    auto dec = [&](int w, float (&v)[4]) {
        v[0] = __builtin_amdgcn_cvt_f32_fp8(w, 0);
        v[1] = __builtin_amdgcn_cvt_f32_fp8(w, 1);
        v[2] = __builtin_amdgcn_cvt_f32_fp8(w, 2);
        v[3] = __builtin_amdgcn_cvt_f32_fp8(w, 3);
    };
    float sum[NS][4];
#pragma unroll
    for (int t = 0; t < NS; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u) sum[t][u] = 0.f;
    float mxr = INFINITY, Mxr = 0.f;
#pragma unroll 4
    for (int64_t row = s0; row < s1; ++row) {
        mxr = fminf(mxr, rn[row]);
        Mxr = fmaxf(Mxr, rn[row]);
#pragma unroll
        for (int t = 0; t < NS; ++t) {
            const int b = 4 * lane + 256 * t;
            if (b < dp) {
                float v[4];
                dec(*(const int *)(L8 + at(row, b)), v);
#pragma unroll
                for (int u = 0; u < 4; ++u) sum[t][u] += v[u];
            }
        }
    }
    const float inv = 1.f / (float)max<int64_t>(s1 - s0, 1);
    float mq[NS][4];
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        const int b = 4 * lane + 256 * t;
        if (b < dp) {
            float v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = fminf(fmaxf(sum[t][u] * inv, -448.f), 448.f);
            int wq = 0;
            wq = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], wq, false);
            wq = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], wq, true);
            *(int *)(C8 + at(sg, b)) = wq;
            dec(wq, mq[t]);
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) mq[t][u] = 0.f;
        }
    }
    float r2max = 0.f, x2max = 0.f;
#pragma unroll 4
    for (int64_t row = s0; row < s1; ++row) {
        float e2 = 0.f, x2 = 0.f;
#pragma unroll
        for (int t = 0; t < NS; ++t) {
            const int b = 4 * lane + 256 * t;
            if (b < dp) {
                float v[4];
                dec(*(const int *)(L8 + at(row, b)), v);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const float e = v[u] - mq[t][u];
                    e2 = fmaf(e, e, e2);
                    x2 = fmaf(v[u], v[u], x2);
                }
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            e2 += __shfl_xor(e2, off, 64);
            x2 += __shfl_xor(x2, off, 64);
        }
        r2max = fmaxf(r2max, __shfl(e2, 0, 64));
        x2max = fmaxf(x2max, __shfl(x2, 0, 64));
    }
    if (lane == 0) {
        constexpr float up = 1.f + 1.f / 1024.f;
        sm[sg] = make_float4(sqrtf(r2max) * up, sqrtf(x2max) * up, Mxr, mxr);
    }
}

// Builds the device index from device rows X (n x d, contiguous), centroids C (nc x d) and labels.
static void ivfl2_set_device(pn_ivfl2 *h, const float *dX, int64_t n, int64_t d, const float *dC, int64_t nc,
                             const int32_t *hlabels, hipStream_t st) {
    // clusters in order, empties left out; ids ascending inside a cluster
    std::vector<int64_t> cnt(nc, 0);
    for (int64_t i = 0; i < n; ++i) {
        PN_REQUIRE(hlabels[i] >= 0 && hlabels[i] < nc, PN_EINVAL, "label out of range");
        ++cnt[hlabels[i]];
    }
    std::vector<int32_t> keep;
    std::vector<int64_t> base(nc, -1);
    std::vector<int64_t> off(1, 0);
    for (int64_t c = 0; c < nc; ++c)
        if (cnt[c]) {
            base[c] = off.back();
            keep.push_back((int32_t)c);
            off.push_back(off.back() + cnt[c]);
        }
    std::vector<int32_t> ids(n);
    {
        std::vector<int64_t> cur(base);
        for (int64_t i = 0; i < n; ++i) ids[cur[hlabels[i]]++] = (int32_t)i;
    }
    const int32_t nl = (int32_t)keep.size();
    h->n = n;
    h->d = d;
    h->nl = nl;
    int32_t *dl = (int32_t *)h->lids.ensure((size_t)n * 4);
    int64_t *dof = (int64_t *)h->loff.ensure((size_t)(nl + 1) * 8);
    PN_HIP(hipMemcpyAsync(dl, ids.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
    PN_HIP(hipMemcpyAsync(dof, off.data(), (size_t)(nl + 1) * 8, hipMemcpyHostToDevice, st));
    std::vector<int32_t> order(nl);
    for (int32_t j = 0; j < nl; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return off[a + 1] - off[a] > off[b + 1] - off[b]; });
    int32_t *dco = (int32_t *)h->corder.ensure((size_t)nl * 4);
    PN_HIP(hipMemcpyAsync(dco, order.data(), (size_t)nl * 4, hipMemcpyHostToDevice, st));
    std::vector<int32_t> clist(nc, -1);
    for (int32_t j = 0; j < nl; ++j) clist[keep[j]] = j;
    h->nc = (int32_t)nc;
    PN_HIP(hipMemcpyAsync(h->clist.ensure((size_t)nc * 4), clist.data(), (size_t)nc * 4, hipMemcpyHostToDevice, st));
    // rows in cluster order (kept for the exact rerank) -> bf16 split (+ zero rows past the end,
    // see the scan kernel)
    float *Xo = (float *)h->X.ensure((size_t)n * d * 4);
    hipLaunchKernelGGL(gather_rows_f32_kernel, dim3((unsigned)ceil_div(n * d, 256)), dim3(256), 0, st, dX, d, dl, n, Xo);
    std::vector<int64_t> soff;
    std::vector<int32_t> sbase;
    sub_partition_order(dX, n, d, off, ids, dl, dof, Xo, soff, sbase, st);
    const int64_t nseg = (int64_t)soff.size() - 1;
    h->nseg = nseg;
    PN_HIP(hipMemcpyAsync(h->soff.ensure((size_t)(nseg + 1) * 8), soff.data(), (size_t)(nseg + 1) * 8,
                          hipMemcpyHostToDevice, st));
    PN_HIP(hipMemcpyAsync(h->sbase.ensure((size_t)(nl + 1) * 4), sbase.data(), (size_t)(nl + 1) * 4,
                          hipMemcpyHostToDevice, st));
    l2_prepare(h->lrows, Xo, n, d, d, st, kIvfKC); // whole K chunks: the scan has no dim tail
    // coarse index over the kept centroids
    bool scan_ok = false;
    const int64_t re = padded_row_elems(PN_METRIC_L2, d, &scan_ok);
    DevBuf cp;
    float *cpp = (float *)cp.ensure((size_t)nl * re * 4);
    PN_HIP(hipMemsetAsync(cpp, 0, (size_t)nl * re * 4, st));
    for (int32_t j = 0; j < nl; ++j)
        PN_HIP(hipMemcpyAsync(cpp + (int64_t)j * re, dC + (int64_t)keep[j] * d, (size_t)d * 4, hipMemcpyDeviceToDevice,
                              st));
    exact_set_device(&h->cent, cpp, nl, d, st);
    { // fp8 residuals to the list centroids, one power-of-two scale mapping the largest |x - c| to <= 224
        float *lc = (float *)h->lcent.ensure((size_t)nl * d * 4);
        PN_HIP(hipMemcpy2DAsync(lc, (size_t)d * 4, cpp, (size_t)re * 4, (size_t)d * 4, nl, hipMemcpyDeviceToDevice, st));
        DevBuf mx;
        unsigned int *dm = (unsigned int *)mx.ensure(4);
        PN_HIP(hipMemsetAsync(dm, 0, 4, st));
        const dim3 g((unsigned)ceil_div(n, (int64_t)4));
        hipLaunchKernelGGL(residual_kernel<0>, g, dim3(256), 0, st, Xo, n, (int)d, h->lrows.dp, dof, nl, lc, 1.f, dm,
                           (uint8_t *)nullptr, (float *)nullptr);
        unsigned int bits = 0;
        PN_HIP(hipMemcpyAsync(&bits, dm, 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipStreamSynchronize(st));
        float m = 0.f;
        std::memcpy(&m, &bits, 4);
        h->sx8 = m > 0.f ? std::exp2(std::floor(std::log2(224.f / m))) : 1.f;
        const int dp = h->lrows.dp;
        // whole 32-row groups (the scan reads a list's groups whole); the rows past n stay zero
        const int64_t npad = ceil_div(n, (int64_t)32) * 32;
        uint8_t *L8 = (uint8_t *)h->l8.ensure((size_t)npad * dp);
        float *rn = (float *)h->rn8.ensure((size_t)npad * 4);
        PN_HIP(hipMemsetAsync(L8, 0, (size_t)npad * dp, st)); // (the last group's rows interleave by chunk)
        PN_HIP(hipMemsetAsync(rn + n, 0, (size_t)(npad - n) * 4, st));
        hipLaunchKernelGGL(residual_kernel<1>, g, dim3(256), 0, st, Xo, n, (int)d, dp, dof, nl, lc, h->sx8, dm, L8, rn);
        // segment centres and radii (centre rows padded to whole 32-row groups of centres)
        const int64_t nseg = h->nseg, ncpad = ceil_div(nseg, (int64_t)32) * 32;
        uint8_t *C8 = (uint8_t *)h->c8.ensure((size_t)ncpad * dp);
        float4 *sm = (float4 *)h->smeta.ensure((size_t)nseg * 16);
        PN_HIP(hipMemsetAsync(C8, 0, (size_t)ncpad * dp, st));
        const int64_t *so = h->soff.as<int64_t>();
        const dim3 gs((unsigned)ceil_div(nseg, (int64_t)4));
        if (dp <= 256)
            hipLaunchKernelGGL(seg_meta_kernel<1>, gs, dim3(256), 0, st, L8, rn, so, nseg, dp, C8, sm);
        else if (dp <= 512)
            hipLaunchKernelGGL(seg_meta_kernel<2>, gs, dim3(256), 0, st, L8, rn, so, nseg, dp, C8, sm);
        else if (dp <= 768)
            hipLaunchKernelGGL(seg_meta_kernel<3>, gs, dim3(256), 0, st, L8, rn, so, nseg, dp, C8, sm);
        else
            hipLaunchKernelGGL(seg_meta_kernel<5>, gs, dim3(256), 0, st, L8, rn, so, nseg, dp, C8, sm);
        PN_HIP(hipGetLastError());
    }
    PN_HIP(hipStreamSynchronize(st)); // host vectors and temporaries are released on return
}

static void ivfl2_lists_enqueue(pn_ivfl2 *h, const float *Q, int64_t re, int64_t nq, int32_t *pr, int np,
                                bool sparse, int k, int prec, int32_t *oi, float *od, int32_t *oc, hipStream_t st);

// Per query, before the list scan: with `swap`, its first probe becomes its nearest probed list
// of at least k2 rows (swapped with slot 0: the bounded fp8 scan takes each query's threshold from
// its first probe, which a shorter list cannot give; the probe set is unchanged and the rerank
// takes a query's candidates by key from all its probes, so the order of a query's probes changes
// no result); qthr (if given) starts at kKey32None (no bound); and the query's first probe (nl for
// none) goes to qkey and is counted in ohist, for the rerank's query order. ohist is zero on entry
// (the rerank kernel clears it after use).
__global__ void query_prep_kernel(int32_t *__restrict__ pr, int64_t nq, int np, const int64_t *__restrict__ loff,
                                  int k2, int swap, uint32_t *__restrict__ qthr, int32_t *__restrict__ qkey,
                                  int32_t *__restrict__ ohist, int nl) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    int32_t *p = pr + q * np;
    if (swap)
        for (int j = 0; j < np; ++j) {
            const int c = p[j];
            if (c >= 0 && loff[c + 1] - loff[c] >= k2) {
                if (j > 0) {
                    p[j] = p[0];
                    p[0] = c;
                }
                break;
            }
        }
    if (qthr) qthr[q] = kKey32None;
    int key = nl;
    for (int j = 0; j < np; ++j)
        if (p[j] >= 0) {
            key = p[j];
            break;
        }
    qkey[q] = key;
    // one atomic per distinct key of the wave (the queries of a hub list share their first list,
    // and per-query atomics on its bin would serialize)
    const int lane = threadIdx.x & 63;
    bool pending = true;
    for (;;) {
        const uint64_t todo = __ballot(pending);
        if (!todo) break;
        const int leader = __builtin_ctzll(todo);
        const int lk = __shfl(key, leader, 64);
        const bool mine = pending && key == lk;
        const uint64_t same = __ballot(mine);
        if (lane == leader) atomicAdd(&ohist[lk], (int)__popcll(same));
        if (mine) pending = false;
    }
}

static void sub_partition_order(const float *dX, int64_t n, int64_t d, const std::vector<int64_t> &off,
                                std::vector<int32_t> &ids, int32_t *dids, const int64_t *dloff, float *Xo,
                                std::vector<int64_t> &soff, std::vector<int32_t> &sbase, hipStream_t st) {
    const int32_t nl = (int32_t)off.size() - 1;
    // segments: the non-empty parts of each partitioned list in chain order, else the whole list
    soff.assign(1, 0);
    sbase.assign(nl + 1, 0);
    auto whole = [&]() {
        soff.assign(1, 0);
        for (int32_t l = 0; l < nl; ++l) {
            soff.push_back(off[l + 1]);
            sbase[l + 1] = l + 1;
        }
    };
    if (d > kSubDims || nl <= 0) {
        whole();
        return;
    }
    std::vector<int32_t> pbase(nl + 1, 0), plist;
    for (int32_t l = 0; l < nl; ++l) {
        const int64_t nr = off[l + 1] - off[l];
        const int32_t m = nr >= 2 * kSubRows ? (int32_t)std::min<int64_t>(kSubMax, nr / kSubRows) : 0;
        pbase[l + 1] = pbase[l] + m;
        for (int32_t j = 0; j < m; ++j) plist.push_back(l);
    }
    const int32_t np_ = pbase[nl];
    if (np_ == 0) {
        whole();
        return;
    }
    DevBuf pb, pl, P, sb;
    int32_t *dpb = (int32_t *)pb.ensure((size_t)(nl + 1) * 4);
    int32_t *dpl = (int32_t *)pl.ensure((size_t)np_ * 4);
    float *dP = (float *)P.ensure((size_t)np_ * d * 4);
    int32_t *dsub = (int32_t *)sb.ensure((size_t)n * 4);
    PN_HIP(hipMemcpyAsync(dpb, pbase.data(), (size_t)(nl + 1) * 4, hipMemcpyHostToDevice, st));
    PN_HIP(hipMemcpyAsync(dpl, plist.data(), (size_t)np_ * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(sub_init_kernel, dim3((unsigned)np_), dim3(256), 0, st, Xo, (int)d, dloff, dpl, dpb, dP);
    const dim3 ga((unsigned)ceil_div(n, (int64_t)4));
    for (int it = 0; it < kSubIters; ++it) {
        hipLaunchKernelGGL(sub_assign_kernel, ga, dim3(256), 0, st, Xo, n, (int)d, dloff, nl, dpb, dP, dsub);
        hipLaunchKernelGGL(sub_update_kernel, dim3((unsigned)np_), dim3(256), 0, st, Xo, (int)d, dloff, dpl, dpb, dsub,
                           dP);
    }
    hipLaunchKernelGGL(sub_assign_kernel, ga, dim3(256), 0, st, Xo, n, (int)d, dloff, nl, dpb, dP, dsub);
    PN_HIP(hipGetLastError());
    std::vector<int32_t> sub(n);
    std::vector<float> hp((size_t)np_ * d);
    PN_HIP(hipMemcpyAsync(sub.data(), dsub, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    PN_HIP(hipMemcpyAsync(hp.data(), dP, hp.size() * 4, hipMemcpyDeviceToHost, st));
    PN_HIP(hipStreamSynchronize(st));
    // Parts in chain order: from part 0, each next part is the nearest unvisited one (lower index
    // on ties), so parts of one clump of the list sit next to each other and the group that
    // straddles two of them stays compact. Then a stable counting sort of the list's positions by
    // the part's place in the chain.
    std::vector<int32_t> nids(ids);
    for (int32_t l = 0; l < nl; ++l) {
        const int32_t m = pbase[l + 1] - pbase[l];
        if (m < 2) {
            soff.push_back(off[l + 1]);
            sbase[l + 1] = (int32_t)soff.size() - 1;
            continue;
        }
        const float *pc = hp.data() + (size_t)pbase[l] * d;
        std::vector<int32_t> place(m, -1);
        int32_t cur_p = 0;
        place[0] = 0;
        for (int32_t t = 1; t < m; ++t) {
            int32_t best = -1;
            double bd = 0.0;
            for (int32_t j = 0; j < m; ++j) {
                if (place[j] >= 0) continue;
                double dd = 0.0;
                for (int64_t e = 0; e < d; ++e) {
                    const double v = (double)pc[(size_t)cur_p * d + e] - (double)pc[(size_t)j * d + e];
                    dd += v * v;
                }
                if (best < 0 || dd < bd) {
                    best = j;
                    bd = dd;
                }
            }
            place[best] = t;
            cur_p = best;
        }
        std::vector<int64_t> cur(m + 1, 0);
        for (int64_t r = off[l]; r < off[l + 1]; ++r) ++cur[place[sub[r]] + 1];
        for (int32_t j = 0; j < m; ++j) cur[j + 1] += cur[j];
        for (int32_t j = 0; j < m; ++j)
            if (cur[j + 1] > cur[j]) soff.push_back(off[l] + cur[j + 1]);
        sbase[l + 1] = (int32_t)soff.size() - 1;
        for (int64_t r = off[l]; r < off[l + 1]; ++r) nids[off[l] + cur[place[sub[r]]]++] = ids[r];
    }
    ids.swap(nids);
    PN_HIP(hipMemcpyAsync(dids, ids.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(gather_rows_f32_kernel, dim3((unsigned)ceil_div(n * d, 256)), dim3(256), 0, st, dX, d, dids, n,
                       Xo);
    PN_HIP(hipGetLastError());
    PN_HIP(hipStreamSynchronize(st)); // (host vectors and temporaries are released on return)
}


// ---------------------------------------------------------------- k > 64: every probed row
// For k above the candidate pipeline's 64 (forest.py:84-105 has no limit), every row of every
// probed list is scored in fp32 (row_sq_dist16, the rerank's arithmetic, so the distances are the
// ones smaller k returns) and each query's keys (distance bits << 32 | id) are sorted; the first
// k are the result, ascending (distance, id).

// rows of the probed lists of query q (probes of -1 count nothing)
__global__ void exh_count_kernel(const int32_t *__restrict__ pr, int64_t nq, int np,
                                 const int64_t *__restrict__ loff, int64_t *__restrict__ rows) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    int64_t t = 0;
    for (int p = 0; p < np; ++p) {
        const int c = pr[q * np + p];
        if (c >= 0) t += loff[c + 1] - loff[c];
    }
    rows[q] = t;
}

// One workgroup per (query, probe) of queries q0 .. q0 + nqc - 1: keys of the list's rows at
// koff[q - q0] + (rows of the query's earlier probes); four rows per wave per round.
template <bool VEC>
__global__ __launch_bounds__(256) void exh_keys_kernel(const float *__restrict__ X, int64_t d,
                                                       const int32_t *__restrict__ lids, const float *__restrict__ Q,
                                                       int64_t qstride, const int32_t *__restrict__ pr, int np,
                                                       const int64_t *__restrict__ loff, int64_t q0,
                                                       const int64_t *__restrict__ koff, uint64_t *__restrict__ keys) {
    __shared__ int64_t s_base;
    const int64_t pair = blockIdx.x;
    const int64_t q = q0 + pair / np;
    const int p = (int)(pair % np);
    const int c = pr[q * np + p];
    if (c < 0) return; // (whole workgroup)
    if (threadIdx.x == 0) {
        int64_t b = koff[q - q0];
        for (int pp = 0; pp < p; ++pp) {
            const int cc = pr[q * np + pp];
            if (cc >= 0) b += loff[cc + 1] - loff[cc];
        }
        s_base = b;
    }
    __syncthreads();
    const int64_t r0 = loff[c], nr = loff[c + 1] - r0, base = s_base;
    const int lane = threadIdx.x & 63, g = lane >> 4, cl = lane & 15, w = threadIdx.x >> 6;
    const float *qrow = Q + q * qstride;
    for (int64_t j0 = 16 * w; j0 < nr; j0 += 64) { // (wave-uniform)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t j = j0 + 4 * u + g;
            const float v = row_sq_dist16<VEC>(qrow, j < nr ? X + (r0 + j) * d : nullptr, d, cl);
            if (cl == 0 && j < nr) keys[base + j] = ((uint64_t)f32_key(v) << 32) | (uint32_t)lids[r0 + j];
        }
    }
}

__global__ void exh_out_kernel(const uint64_t *__restrict__ sorted, const int64_t *__restrict__ koff, int64_t q0,
                               int64_t nqc, int k, int32_t *__restrict__ oi, float *__restrict__ od,
                               int32_t *__restrict__ oc) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nqc * k) return;
    const int64_t qi = t / k, j = t - qi * k;
    const int64_t n = koff[qi + 1] - koff[qi];
    const int64_t o = (q0 + qi) * k + j;
    if (j < n) {
        const uint64_t key = sorted[koff[qi] + j];
        oi[o] = (int32_t)(uint32_t)(key & 0xFFFFFFFFull);
        od[o] = key_f32((uint32_t)(key >> 32));
    } else {
        oi[o] = -1;
        od[o] = 0.f;
    }
    if (j == 0) oc[q0 + qi] = (int32_t)min<int64_t>(n, k);
}

// Host-synchronous (the row counts size the key buffers); queries in chunks of at most
// kExhKeys keys (or one query).
constexpr int64_t kExhKeys = int64_t(1) << 26;
static void ivfl2_exhaustive(pn_ivfl2 *h, const float *Q, int64_t re, int64_t nq, const int32_t *pr, int np, int k,
                             int32_t *oi, float *od, int32_t *oc, hipStream_t st) {
    StageScope sc(st, STAGE_RERANK);
    const int64_t d = h->d;
    DevBuf rows_d, koff_d, keys_a, keys_b, tmp;
    int64_t *rows = (int64_t *)rows_d.ensure((size_t)nq * 8);
    hipLaunchKernelGGL(exh_count_kernel, dim3((unsigned)ceil_div(nq, (int64_t)256)), dim3(256), 0, st, pr, nq, np,
                       h->loff.as<int64_t>(), rows);
    std::vector<int64_t> hr(nq);
    PN_HIP(hipMemcpyAsync(hr.data(), rows, (size_t)nq * 8, hipMemcpyDeviceToHost, st));
    PN_HIP(hipStreamSynchronize(st));
    const bool vec = d % 4 == 0 && re % 4 == 0;
    int64_t qa = 0;
    while (qa < nq) {
        int64_t qb = qa, tot = 0;
        while (qb < nq && (qb == qa || tot + hr[qb] <= kExhKeys)) tot += hr[qb++];
        PN_REQUIRE(tot < (int64_t)INT32_MAX, PN_EINVAL, "IVF search: too many probed rows for one query");
        const int64_t nqc = qb - qa;
        std::vector<int64_t> ko(nqc + 1, 0);
        for (int64_t i = 0; i < nqc; ++i) ko[i + 1] = ko[i] + hr[qa + i];
        int64_t *koff = (int64_t *)koff_d.ensure((size_t)(nqc + 1) * 8);
        PN_HIP(hipMemcpyAsync(koff, ko.data(), (size_t)(nqc + 1) * 8, hipMemcpyHostToDevice, st));
        uint64_t *ka = (uint64_t *)keys_a.ensure((size_t)std::max<int64_t>(tot, 1) * 8);
        uint64_t *kb = (uint64_t *)keys_b.ensure((size_t)std::max<int64_t>(tot, 1) * 8);
        if (tot > 0) {
            const dim3 gk((unsigned)(nqc * np));
            if (vec)
                hipLaunchKernelGGL(exh_keys_kernel<true>, gk, dim3(256), 0, st, h->X.as<float>(), d,
                                   h->lids.as<int32_t>(), Q, re, pr, np, h->loff.as<int64_t>(), qa, koff, ka);
            else
                hipLaunchKernelGGL(exh_keys_kernel<false>, gk, dim3(256), 0, st, h->X.as<float>(), d,
                                   h->lids.as<int32_t>(), Q, re, pr, np, h->loff.as<int64_t>(), qa, koff, ka);
            size_t tb = 0;
            PN_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tb, ka, kb, (int)tot, (int)nqc, koff, koff + 1,
                                                              0, 64, st));
            void *tp = tmp.ensure(tb);
            PN_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(tp, tb, ka, kb, (int)tot, (int)nqc, koff, koff + 1, 0, 64,
                                                              st));
        }
        hipLaunchKernelGGL(exh_out_kernel, dim3((unsigned)ceil_div(nqc * k, (int64_t)256)), dim3(256), 0, st, kb, koff,
                           qa, nqc, k, oi, od, oc);
        PN_HIP(hipGetLastError());
        PN_HIP(hipStreamSynchronize(st)); // (the chunk's host offsets and buffers are reused)
        qa = qb;
    }
}

// ABI precision -> scan variant: 0 bf16 (1 product), 1 bf16 hi/lo (3 products), 2 fp8 e4m3
static int prec_of(int32_t precision) {
    PN_REQUIRE(precision >= 0 && precision <= 2, PN_EINVAL, "precision must be 0 (bf16), 1 (bf16x3) or 2 (fp8)");
    return precision == 1 ? 3 : (precision == 2 ? 8 : 1);
}

static void ivfl2_search_enqueue(pn_ivfl2 *h, const float *dq, int64_t qstride_in, int64_t nq, int nprobe, int k,
                                 int prec, int32_t *oi, float *od, int32_t *oc, hipStream_t st) {
    const int64_t d = h->d;
    const int np = std::min<int>(nprobe, h->nl);
    bool scan_ok = false;
    const int64_t re = padded_row_elems(PN_METRIC_L2, d, &scan_ok);
    // queries with the centroid index's padded stride
    const float *Q = dq;
    if (qstride_in != re) {
        float *qp = (float *)h->qpad.ensure((size_t)nq * re * 4);
        PN_HIP(hipMemsetAsync(qp, 0, (size_t)nq * re * 4, st));
        PN_HIP(hipMemcpy2DAsync(qp, (size_t)re * 4, dq, (size_t)qstride_in * 4, (size_t)d * 4, nq,
                                hipMemcpyDeviceToDevice, st));
        Q = qp;
    }
    if (np <= 0) {
        PN_HIP(hipMemsetAsync(oi, 0xFF, (size_t)nq * k * 4, st));
        PN_HIP(hipMemsetAsync(oc, 0, (size_t)nq * 4, st));
        return;
    }
    int32_t *pr = (int32_t *)h->probes.ensure((size_t)nq * np * 4);
    void *pd = h->pdist.ensure((size_t)nq * np * 4);
    {
        StageScope sc(st, STAGE_COARSE);
        exact_search_enqueue(&h->cent, Q, nq, np, pr, pd, st);
    }
    ivfl2_lists_enqueue(h, Q, re, nq, pr, np, false, k, prec, oi, od, oc, st);
}

// The list scan and rerank for given probes pr[nq x np] (list indices; with `sparse`, -1 marks an
// empty slot). Q has row stride re.
// pr: the handle's own probe buffer (list indices), reordered per query by query_prep_kernel for
// the bounded fp8 scans.
static void ivfl2_lists_enqueue(pn_ivfl2 *h, const float *Q, int64_t re, int64_t nq, int32_t *pr, int np,
                                bool sparse, int k, int prec, int32_t *oi, float *od, int32_t *oc, hipStream_t st) {
    if (k > kIvfK2) { // (host entry point only: the device ones take k <= 64)
        ivfl2_exhaustive(h, Q, re, nq, pr, np, k, oi, od, oc, st);
        return;
    }
    const int64_t d = h->d;
    // bf16 split of the queries
    const int dp = h->lrows.dp;
    __bf16 *Qh = (__bf16 *)h->qsplit_h.ensure((size_t)nq * dp * 2);
    __bf16 *Ql = (__bf16 *)h->qsplit_l.ensure((size_t)nq * dp * 2);
    float *qn = (float *)h->qn.ensure((size_t)nq * 4);
    const int64_t total = nq * np;
    // per-list pair counts: zero between passes (probe_scatter_kernel clears them after use), so a
    // new allocation is cleared once here
    const bool lc_new = !h->lcount.p || h->lcount.bytes < (size_t)h->nl * 4;
    int32_t *lc = (int32_t *)h->lcount.ensure((size_t)h->nl * 4);
    if (lc_new) PN_HIP(hipMemsetAsync(lc, 0, (size_t)h->nl * 4, st));
    // the rerank's query order: a histogram of first probes (zero between searches: the rerank
    // kernel clears it after use, so a new allocation is cleared once here; a search that stopped on
    // an error before its rerank leaves it dirty: cleared here then), keys and the order
    if (prec == 8) PN_REQUIRE(dp <= 1152, PN_EINVAL, "fp8 precision serves d up to 1152");
    const bool oh_new = !h->ohist.p || h->ohist.bytes < (size_t)(h->nl + 1) * 4;
    int32_t *ohist = (int32_t *)h->ohist.ensure((size_t)(h->nl + 1) * 4);
    if (oh_new || h->ohist_dirty) PN_HIP(hipMemsetAsync(ohist, 0, (size_t)(h->nl + 1) * 4, st));
    h->ohist_dirty = true;
    int32_t *qkey = (int32_t *)h->qkey.ensure((size_t)nq * 4);
    int32_t *qorder = (int32_t *)h->qorder.ensure((size_t)nq * 4);
    int32_t *cur = (int32_t *)h->cursor.ensure((size_t)h->nl * 4);
    int32_t *pq = (int32_t *)h->pair_q.ensure((size_t)total * 4);
    int32_t *pp = (int32_t *)h->pair_p.ensure((size_t)total * 4);
    int32_t *pc = (int32_t *)h->pair_c.ensure((size_t)total * 4);
    int32_t *psl = (int32_t *)h->pslot.ensure((size_t)total * 4);
    // queries per item: the fp8 scan takes up to 128 (two query groups share each row fetched),
    // fewer in its bounded second pass (ivfl2.fp8_iq2: most items there need no row at all, and
    // the few that do spread over more workgroups); the bf16 scans 64
    const int iq_b = prec == 8 ? (int)std::min<int64_t>(kIvf8IQ, std::max<int64_t>(32, tune_get("ivfl2.fp8_iq2")))
                               : kIvfIQ;
    // each cluster has at most one partly filled item
    const int64_t max_items = ceil_div(total, (int64_t)std::min(kIvfIQ, iq_b)) + h->nl;
    int4 *items = (int4 *)h->items.ensure((size_t)max_items * 16);
    int32_t *bins = (int32_t *)h->ntot.ensure(kBinsInts * 4);
    int32_t *itoff = (int32_t *)h->itoff.ensure((size_t)h->nl * 4);
    int64_t *woff = (int64_t *)h->woff.ensure((size_t)h->nl * 8);
    const int nranges = tune_get("ivfl2.xcd_ranges") ? kXcds : 1;
    // candidate slots per (query, probe): the bf16 scan merges its lists into the KP = 32 best,
    // the fp8 scan (more lists, coarser keys) into the 64 best
    const int kl = k <= 16 ? 8 : 16;
    const int cap = prec == 8 || k > kIvfKP ? kIvf8Cap : kIvfKP; // (k > 32: 64 per probe)
    uint64_t *cand = (uint64_t *)h->cand.ensure((size_t)total * cap * 8);
    // with per-query probe counts, each query's slots are packed after the previous query's
    // (qoff); slots of unused probes inside a query's span stay empty
    const int32_t *qoff = nullptr;
    {
        StageScope sc(st, STAGE_OTHER);
        if (sparse) {
            int32_t *span = (int32_t *)h->qoff.ensure((size_t)(nq + 1) * 4);
            hipLaunchKernelGGL(probe_span_kernel, dim3((unsigned)ceil_div(nq + 1, (int64_t)256)), dim3(256), 0, st, pr,
                               nq, np, span);
            hipLaunchKernelGGL(order_scan_kernel, dim3(1), dim3(256), 0, st, span, (int)(nq + 1));
            hipLaunchKernelGGL(fill_slots_kernel, dim3(2048), dim3(256), 0, st, cand, span, nq, cap, kKey64None);
            qoff = span;
        }
        if (prec != 8) l2_split_rows(Q, nq, d, re, dp, Qh, Ql, qn, st); // (the fp8 scan converts in LDS)
    }
    // candidates reranked per query: ivfl2.k2 (k' >= 4k is enough for the rerank to recover
    // the exact top-k in practice), at least k, at most kIvfK2 = 64 (one per lane)
    const int k2 = (int)std::min<int64_t>(kIvfK2, std::max<int64_t>({(int64_t)k, 4 * (int64_t)k, tune_get("ivfl2.k2")}));
    // fp8 with segment bounds (the default; d <= 768, the query tile's fragments in registers):
    // each query's threshold from its first probe's segments, then one pass over all probes that
    // skips the row groups no query of a query group can need (ivfl2.prune; the results are those
    // of the same pass without the skipping)
    const int dp8 = h->lrows.dp;
    const bool bounded = np > 1 && prec == 8 && cap >= k2 && dp8 <= 768 && h->nseg > 0 &&
                         tune_get("ivfl2.bound") != 0;
    const bool prune = bounded && tune_get("ivfl2.prune") != 0;
    // else, fp8: optionally two passes, the first probe scanned first for the bound
    const bool k2_pass = !bounded && np > 1 && prec == 8 && cap >= k2 && tune_get("ivfl2.two_pass") != 0;
    uint32_t *qthr = bounded || k2_pass || (np > 1 && k <= kIvfKP && tune_get("ivfl2.pass_margin_pct") > 0)
                         ? (uint32_t *)h->qthr.ensure((size_t)nq * 4)
                         : nullptr;
    {
        StageScope sc(st, STAGE_OTHER);
        hipLaunchKernelGGL(query_prep_kernel, dim3((unsigned)ceil_div(nq, (int64_t)256)), dim3(256), 0, st, pr, nq, np,
                           h->loff.as<int64_t>(), k2, (bounded || k2_pass) ? 1 : 0, bounded ? qthr : nullptr, qkey,
                           ohist, h->nl);
    }
    // pairs of probe slots p_lo .. p_lo + npw - 1 -> items of at most iq pairs of one cluster; the
    // first call also finishes the rerank's query order
    bool first_bucket = true;
    auto bucket = [&](int p_lo, int npw) {
        StageScope sc(st, STAGE_OTHER);
        const int64_t tw = nq * npw;
        const unsigned pchunks = (unsigned)ceil_div(tw, (int64_t)kPairChunk);
        hipLaunchKernelGGL(probe_count_kernel, dim3(pchunks), dim3(256), 0, st, pr, tw, np, p_lo, npw, h->nl, lc);
        const int iq = prec == 8 ? (p_lo > 0 ? iq_b : kIvf8IQ) : kIvfIQ; // queries per item
        hipLaunchKernelGGL(build_items_kernel, dim3(1), dim3(256), 0, st, lc, h->corder.as<int32_t>(),
                           h->loff.as<int64_t>(), h->nl, cur, itoff, woff, bins, nranges, iq,
                           first_bucket ? ohist : nullptr, h->nl + 1);
        const int eblocks = (int)ceil_div(max_items, (int64_t)256);
        const int oblocks = first_bucket ? (int)ceil_div(nq, (int64_t)256) : 0;
        hipLaunchKernelGGL(expand_items_kernel, dim3((unsigned)(eblocks + oblocks)), dim3(256), 0, st, lc,
                           h->corder.as<int32_t>(), h->loff.as<int64_t>(), h->nl, cur, itoff, woff, max_items, nranges,
                           iq, items, bins, eblocks, qkey, nq, ohist, qorder);
        first_bucket = false;
        hipLaunchKernelGGL(probe_scatter_kernel, dim3(pchunks), dim3(256), 0, st, pr, tw, np, p_lo, npw, h->nl, cur, pq,
                           pp, pc, psl, lc);
        if (prec == 8) { // fp8 query residuals of every pair
            uint8_t *R8 = (uint8_t *)h->qres8.ensure((size_t)total * dp);
            float *Rn = (float *)h->qresn.ensure((size_t)total * 4);
            // norms of the fp8 residuals for the segment bounds
            float *Rq = bounded ? (float *)h->qres8n.ensure((size_t)total * 4) : nullptr;
            const float *lcp = h->lcent.as<float>();
            // rows of whole 256-dim steps (d = dp = 256, 512 or 768, float4-aligned): no masking
            const int nvr = dp <= 256 ? 1 : (dp <= 512 ? 2 : (dp <= 768 ? 3 : 5));
            const bool full = d == dp && dp == 256 * nvr && re % 4 == 0 && ((uintptr_t)Q & 15) == 0;
            // many probes per query (a learned probe's): the XCD split of the centroid rows
            const int64_t xnp = tune_get("ivfl2.residual_xcd_np");
            const bool xsplit = xnp > 0 && npw >= xnp;
            const dim3 gq((unsigned)(ceil_div(nq, (int64_t)4) * (xsplit ? kXcds : 1)));
            auto go = [&](auto nv, auto fl) {
                hipLaunchKernelGGL((query_residual_kernel<decltype(nv)::value, decltype(fl)::value>), gq, dim3(256), 0,
                                   st, Q, re, (int)d, dp, lcp, h->sx8, pr, nq, np, p_lo, npw, psl, R8, Rn, Rq,
                                   xsplit);
            };
            auto by_nv = [&](auto fl) {
                if (nvr == 1) go(std::integral_constant<int, 1>{}, fl);
                else if (nvr == 2) go(std::integral_constant<int, 2>{}, fl);
                else if (nvr == 3) go(std::integral_constant<int, 3>{}, fl);
                else go(std::integral_constant<int, 5>{}, std::false_type{});
            };
            if (full) by_nv(std::true_type{});
            else by_nv(std::false_type{});
        }
        PN_HIP(hipGetLastError());
    };
    const int64_t any_at = max_items * 4 * kPruneWords; // (the "any" words after the group words)
    // segment bounds over the work items: MODE 0 thresholds into qthr, MODE 1 group masks
    auto seg_bounds = [&](int mode, uint32_t *qthr, uint32_t *pa, int64_t any_at) {
        auto go = [&](auto ncc, auto mc) {
            hipLaunchKernelGGL((seg_bound_kernel<decltype(ncc)::value, decltype(mc)::value>), dim3((unsigned)max_items),
                               dim3(256), 0, st, h->c8.as<uint8_t>(), h->smeta.as<float4>(), h->soff.as<int64_t>(),
                               h->sbase.as<int32_t>(), h->loff.as<int64_t>(), dp8, h->sx8, h->qres8.as<uint8_t>(),
                               h->qresn.as<float>(), h->qres8n.as<float>(), items, bins, pq, pp, qthr, pa, any_at, k2);
        };
        auto by_mode = [&](auto ncc) {
            if (mode == 0)
                go(ncc, std::integral_constant<int, 0>{});
            else
                go(ncc, std::integral_constant<int, 1>{});
        };
        if (dp8 <= 256)
            by_mode(std::integral_constant<int, 4>{});
        else if (dp8 <= 512)
            by_mode(std::integral_constant<int, 8>{});
        else
            by_mode(std::integral_constant<int, 12>{});
        PN_HIP(hipGetLastError());
    };
    auto scan = [&](const uint32_t *qthr) {
        StageScope sc(st, STAGE_SCAN);
        // a persistent grid of wg_per_cu workgroups per CU (a multiple of kXcds) drains the item ranges
        const int64_t per_xcd = std::max<int64_t>(1, ceil_div((int64_t)h->cus, (int64_t)kXcds) *
                                                         std::max<int64_t>(1, tune_get("ivfl2.wg_per_cu")));
        const dim3 grid((unsigned)std::min<int64_t>(per_xcd, max_items) * kXcds);
        auto launch = [&](auto pc, auto kc) {
            hipLaunchKernelGGL((ivfl2_scan_kernel<decltype(pc)::value, decltype(kc)::value>), grid, dim3(256), 0, st,
                               h->lrows.hi.as<__bf16>(), h->lrows.lo.as<__bf16>(), h->lrows.norm.as<float>(),
                               h->loff.as<int64_t>(), dp, Qh, Ql, qn, items, bins, nranges, pq, pp, np, qoff, qthr,
                               cand, cap);
        };
        using P1 = std::integral_constant<int, 1>;
        using P3 = std::integral_constant<int, 3>;
        using K8 = std::integral_constant<int, 8>;
        using K16 = std::integral_constant<int, 16>;
        auto launch8 = [&](auto kc) {
            constexpr int KLV = decltype(kc)::value;
            // LDS: the residual block, then the row norms of lists of up to nmax rows
            constexpr size_t kLdsMax = 148 * 1024;
            const size_t rblk = (size_t)kIvf8IQ * (dp + 16);
            PN_REQUIRE(rblk <= kLdsMax, PN_EINVAL, "fp8 precision serves d up to 1152");
            const int nmax = (int)std::min<size_t>(8192, (kLdsMax - rblk) / 4 / 512 * 512); // (>= 512 at dp <= 1152)
            PN_REQUIRE(nmax >= 512, PN_EINVAL, "fp8 precision serves d up to 1152");
            const size_t lds = std::max(rblk + (size_t)nmax * 4, (size_t)16 * 64 * KLV * 8); // (and the lists)
            const dim3 grid8((unsigned)std::min<int64_t>(ceil_div((int64_t)h->cus, (int64_t)kXcds) *
                                                             std::max<int64_t>(1, tune_get("ivfl2.fp8_wg_per_cu")),
                                                         max_items) * kXcds);
            // group pruning of the bounded pass (the masks written by seg_bound_kernel, mode 1)
            const uint32_t *act = prune && qthr ? h->pact.as<uint32_t>() : nullptr;
            unsigned long long *wk = (unsigned long long *)h->work.ensure((size_t)grid8.x * 16);
            h->work_n = (int64_t)grid8.x;
            auto go = [&](auto pdc) {
                constexpr int PDV = decltype(pdc)::value;
                if (lds > 64 * 1024)
                    PN_HIP(hipFuncSetAttribute((const void *)ivfl2_scan_fp8_kernel<KLV, PDV>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                hipLaunchKernelGGL((ivfl2_scan_fp8_kernel<KLV, PDV>), grid8, dim3(512), lds, st, h->l8.as<uint8_t>(),
                                   h->rn8.as<float>(), h->loff.as<int64_t>(), dp, h->sx8, h->qres8.as<uint8_t>(),
                                   h->qresn.as<float>(), items, bins, nranges, pq, pp, np, qoff, qthr, cand, nmax, act,
                                   any_at, wk);
            };
            if (tune_get("ivfl2.fp8_ahead") == 3)
                go(std::integral_constant<int, 3>{});
            else
                go(std::integral_constant<int, 2>{});
        };
        if (prec == 8)
            kl == 8 ? launch8(K8{}) : launch8(K16{});
        else if (prec == 3)
            k <= 16 ? launch(P3{}, K8{}) : launch(P3{}, K16{});
        else
            k <= 16 ? launch(P1{}, K8{}) : launch(P1{}, K16{});
        PN_HIP(hipGetLastError());
    };
    // Two passes when a bound helps: the first probe of every query, then the others with a bound
    // from the first. fp8 (cap = K2 slots per pair; the default): the K2-th candidate of the first
    // probe, which no candidate the rerank takes can exceed (pass_bound_k2_kernel); optionally
    // (pass_margin_pct > 0) the k-th approximation times a margin; otherwise one pass.
    const int64_t margin_pct = tune_get("ivfl2.pass_margin_pct");
    if (bounded) {
        bucket(0, np);
        {
            StageScope sc(st, STAGE_OTHER);
            // group words per (item, query tile), then one "any group needed" word per (item, tile)
            uint32_t *pa = (uint32_t *)h->pact.ensure((size_t)(any_at + max_items * 4) * 4);
            seg_bounds(0, qthr, pa, any_at);
            if (prune) seg_bounds(1, qthr, pa, any_at);
        }
        scan(qthr);
    } else if (k2_pass || (np > 1 && k <= kIvfKP && margin_pct > 0)) {
        bucket(0, 1);
        scan(nullptr);
        {
            StageScope sc(st, STAGE_OTHER);
            const dim3 g((unsigned)ceil_div(nq, (int64_t)4));
            if (k2_pass)
                hipLaunchKernelGGL(pass_bound_k2_kernel, g, dim3(256), 0, st, cand, nq, np, cap, qoff, k2, qthr);
            else
                hipLaunchKernelGGL(pass_bound_kernel, g, dim3(256), 0, st, cand, nq, np, cap, qoff, k,
                                   1.f + (float)margin_pct / 100.f, qthr);
        }
        bucket(1, np - 1);
        scan(qthr);
    } else {
        bucket(0, np);
        scan(nullptr);
    }
    {
        StageScope sc(st, STAGE_RERANK);
        const bool vec = d % 4 == 0 && re % 4 == 0;
        // queries by first probed list (qorder, from the first bucketing)
        const int64_t per_xcd = ceil_div(ceil_div(nq, (int64_t)4), (int64_t)kXcds);
        dim3 g((unsigned)(per_xcd * kXcds));
        auto launch = [&](auto v) {
            hipLaunchKernelGGL((ivfl2_rerank_kernel<decltype(v)::value>), g, dim3(256), 0, st, h->X.as<float>(), d,
                               h->lids.as<int32_t>(), Q, re, cand, np, cap, qoff, nq, k, oi, od, oc, qorder, per_xcd,
                               k2, ohist, h->nl + 1);
        };
        vec ? launch(std::true_type{}) : launch(std::false_type{});
        PN_HIP(hipGetLastError());
        h->ohist_dirty = false;
    }
}

} // namespace pn

using namespace pn;

extern "C" pn_status pn_ivfl2_create(pn_ivfl2 **out) {
    return guard([&] {
        PN_REQUIRE(out, PN_EINVAL, "out is null");
        auto *h = new pn_ivfl2();
        h->device = current_device();
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) == hipSuccess && cus > 0)
            h->cus = cus;
        exact_init(&h->cent, PN_METRIC_L2);
        *out = h;
    });
}

extern "C" void pn_ivfl2_destroy(pn_ivfl2 *h) {
    if (!h) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(h->device);
    delete h;
    if (prev >= 0) (void)hipSetDevice(prev);
}

extern "C" pn_status pn_ivfl2_build(pn_ivfl2 *h, const float *rows, int64_t n, int64_t d, int64_t n_clusters,
                                    int64_t max_iter, uint32_t seed, int32_t *labels_out, float *centroids_out) {
    return guard([&] {
        PN_REQUIRE(h && rows && labels_out && centroids_out, PN_EINVAL, "null pointer");
        PN_REQUIRE(n > 0 && d > 0, PN_EINVAL, "data must be a non-empty 2-D array");
        PN_REQUIRE(n_clusters >= 1 && n_clusters <= n, PN_EINVAL, "n_clusters must be in [1, N]");
        PN_REQUIRE(n < (int64_t)INT32_MAX, PN_EINVAL, "too many rows");
        DeviceScope ds(h->device);
        hipStream_t st = lib_stream();
        DevBuf X, L, C;
        float *dX = (float *)X.ensure((size_t)n * d * 4);
        int32_t *dL = (int32_t *)L.ensure((size_t)n * 4);
        float *dC = (float *)C.ensure((size_t)n_clusters * d * 4);
        PN_HIP(hipMemcpyAsync(dX, rows, (size_t)n * d * 4, hipMemcpyHostToDevice, st));
        kmeans_l2_device(dX, n, d, d, n_clusters, max_iter, seed, dL, dC, st);
        PN_HIP(hipMemcpyAsync(labels_out, dL, (size_t)n * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipMemcpyAsync(centroids_out, dC, (size_t)n_clusters * d * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipStreamSynchronize(st));
        ivfl2_set_device(h, dX, n, d, dC, n_clusters, labels_out, st);
    });
}

extern "C" pn_status pn_ivfl2_set(pn_ivfl2 *h, const float *rows, int64_t n, int64_t d, const float *centroids,
                                  int64_t nc, const int32_t *labels) {
    return guard([&] {
        PN_REQUIRE(h && rows && centroids && labels, PN_EINVAL, "null pointer");
        PN_REQUIRE(n > 0 && d > 0 && nc > 0, PN_EINVAL, "empty index data");
        PN_REQUIRE(n < (int64_t)INT32_MAX, PN_EINVAL, "too many rows");
        DeviceScope ds(h->device);
        hipStream_t st = lib_stream();
        DevBuf X, C;
        float *dX = (float *)X.ensure((size_t)n * d * 4);
        float *dC = (float *)C.ensure((size_t)nc * d * 4);
        PN_HIP(hipMemcpyAsync(dX, rows, (size_t)n * d * 4, hipMemcpyHostToDevice, st));
        PN_HIP(hipMemcpyAsync(dC, centroids, (size_t)nc * d * 4, hipMemcpyHostToDevice, st));
        ivfl2_set_device(h, dX, n, d, dC, nc, labels, st);
    });
}

extern "C" pn_status pn_ivfl2_scan_work(pn_ivfl2 *h, int64_t *mfma_ops, int64_t *row_bytes) {
    return guard([&] {
        PN_REQUIRE(h && mfma_ops && row_bytes, PN_EINVAL, "null pointer");
        DeviceScope ds(h->device);
        std::vector<uint64_t> v((size_t)h->work_n * 2, 0);
        if (h->work_n) PN_HIP(hipMemcpy(v.data(), h->work.p, v.size() * 8, hipMemcpyDeviceToHost));
        int64_t m = 0, b = 0;
        for (int64_t i = 0; i < h->work_n; ++i) {
            m += (int64_t)v[2 * i];
            b += (int64_t)v[2 * i + 1];
        }
        *mfma_ops = m;
        *row_bytes = b;
    });
}

extern "C" pn_status pn_ivfl2_info(const pn_ivfl2 *h, int64_t *n, int64_t *d, int32_t *n_lists) {
    return guard([&] {
        PN_REQUIRE(h && n && d && n_lists, PN_EINVAL, "null argument");
        *n = h->n;
        *d = h->d;
        *n_lists = h->nl;
    });
}

extern "C" pn_status pn_ivfl2_search(pn_ivfl2 *h, const float *queries, int64_t nq, int64_t d, int32_t nprobe,
                                     int32_t k, int32_t precision, int64_t *out_idx, float *out_dist,
                                     int32_t *out_count) {
    return guard([&] {
        PN_REQUIRE(h, PN_EINVAL, "index is null");
        PN_REQUIRE(h->n > 0, PN_ERUNTIME, "Index is empty - call set() first");
        PN_REQUIRE(nq >= 0 && k >= 0, PN_EINVAL, "negative size");
        PN_REQUIRE(d == h->d, PN_EINVAL, "query dimension differs from the index dimension");
        if (nq == 0 || k == 0) {
            if (nq && out_count) std::memset(out_count, 0, (size_t)nq * 4);
            return;
        }
        PN_REQUIRE(queries && out_idx && out_dist && out_count, PN_EINVAL, "null pointer");
        DeviceScope ds(h->device);
        hipStream_t st = lib_stream();
        DevBuf dq, oi, od, oc;
        float *q = (float *)dq.ensure((size_t)nq * d * 4);
        PN_HIP(hipMemcpyAsync(q, queries, (size_t)nq * d * 4, hipMemcpyHostToDevice, st));
        int32_t *i32 = (int32_t *)oi.ensure((size_t)nq * k * 4);
        float *f32 = (float *)od.ensure((size_t)nq * k * 4);
        int32_t *c32 = (int32_t *)oc.ensure((size_t)nq * 4);
        ivfl2_search_enqueue(h, q, d, nq, nprobe, k, prec_of(precision), i32, f32, c32, st);
        std::vector<int32_t> hi((size_t)nq * k);
        PN_HIP(hipMemcpyAsync(hi.data(), i32, hi.size() * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipMemcpyAsync(out_dist, f32, (size_t)nq * k * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipMemcpyAsync(out_count, c32, (size_t)nq * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipStreamSynchronize(st));
        for (size_t i = 0; i < hi.size(); ++i) out_idx[i] = hi[i];
    });
}

extern "C" pn_status pn_ivfl2_search_device(pn_ivfl2 *h, const float *queries, int64_t nq, int32_t nprobe, int32_t k,
                                            int32_t precision, int32_t *out_idx, float *out_dist, int32_t *out_count,
                                            void *stream) {
    return guard([&] {
        PN_REQUIRE(h, PN_EINVAL, "index is null");
        PN_REQUIRE(h->n > 0, PN_ERUNTIME, "Index is empty - call set() first");
        PN_REQUIRE(nq >= 0 && k > 0 && k <= 64, PN_EINVAL, "bad size");
        if (nq == 0) return;
        DeviceScope ds(h->device);
        ivfl2_search_enqueue(h, queries, h->d, nq, nprobe, k, prec_of(precision), out_idx, out_dist, out_count,
                             (hipStream_t)stream);
    });
}

extern "C" pn_status pn_ivfl2_search_probes_device(pn_ivfl2 *h, const float *queries, int64_t nq,
                                                   const int32_t *probes, const int32_t *counts, int32_t np_max,
                                                   int32_t k, int32_t precision, int32_t *out_idx, float *out_dist,
                                                   int32_t *out_count, void *stream) {
    return guard([&] {
        PN_REQUIRE(h, PN_EINVAL, "index is null");
        PN_REQUIRE(h->n > 0, PN_ERUNTIME, "Index is empty - call set() first");
        PN_REQUIRE(nq >= 0 && k > 0 && k <= 64 && np_max > 0, PN_EINVAL, "bad size");
        if (nq == 0) return;
        PN_REQUIRE(queries && probes && counts && out_idx && out_dist && out_count, PN_EINVAL, "null pointer");
        DeviceScope ds(h->device);
        hipStream_t st = (hipStream_t)stream;
        bool scan_ok = false;
        const int64_t re = padded_row_elems(PN_METRIC_L2, h->d, &scan_ok);
        const float *Q = queries;
        if (re != h->d) {
            float *qp = (float *)h->qpad.ensure((size_t)nq * re * 4);
            PN_HIP(hipMemsetAsync(qp, 0, (size_t)nq * re * 4, st));
            PN_HIP(hipMemcpy2DAsync(qp, (size_t)re * 4, queries, (size_t)h->d * 4, (size_t)h->d * 4, nq,
                                    hipMemcpyDeviceToDevice, st));
            Q = qp;
        }
        int32_t *pm = (int32_t *)h->pmap.ensure((size_t)nq * np_max * 4);
        hipLaunchKernelGGL(map_probes_kernel, dim3((unsigned)ceil_div(nq * np_max, (int64_t)256)), dim3(256), 0, st,
                           probes, counts, nq, np_max, h->clist.as<int32_t>(), h->nc, pm);
        ivfl2_lists_enqueue(h, Q, re, nq, pm, np_max, true, k, prec_of(precision), out_idx, out_dist, out_count,
                            st);
    });
}
