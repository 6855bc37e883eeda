// pn_ivfl2: the search of IVFFlatL2Index (pynear/forest.py:63-115) on the GPU.
//
// The reference loops over queries and probed clusters in Python: a GEMM for the coarse
// distances (:76, :200-204), argpartition to nprobe (:77-80), a GEMV per (query, cluster)
// (:189-194), a per-cluster argpartition and a global heapq (:84-105), ascending squared L2.
// Here one search is four device stages:
//   1. coarse: exact L2 top-nprobe clusters per query (the exact search over the centroids).
//   2. bucketing: (query, probe) pairs are grouped by cluster and cut into work items of at most
//      32 queries of one cluster, so every probed cluster is streamed once per search.
//   3. scan: per item, MFMA tiles of (32 cluster rows x 32 queries) in the GEMM form
//      |q|^2 + |x|^2 - 2 q.x on a bf16 copy of the rows (hi only, or hi+lo for 3 products);
//      each (query, probe) keeps its KP best candidates by the approximation.
//   4. rerank: per query the K2 = 64 best candidates over all probes are recomputed exactly
//      (squared L2 in the reference kernel's operation order, fp32 rows) and the k smallest
//      (distance, id) are returned.
// Clusters are stored contiguously in cluster order (ids ascending within a cluster); empty
// clusters are left out, as the reference's per-cluster arrays leave them out (forest.py:51-59).
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "device_util.hpp"
#include "exact_index.hpp"
#include "kernels.hpp"
#include "kmeans_l2.hpp"
#include "l2_mfma.hpp"
#include "ref_arith.hpp"

namespace pn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kIvfKL = 16; // per-lane candidate list
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
    pn::DevBuf qpad, qsplit_h, qsplit_l, qn, probes, pdist, lcount, loffq, cursor, pair_q, pair_p;
    pn::DevBuf items, ntot, cand, oidx, odist, ocnt;
};

namespace pn {

__global__ void probe_count_kernel(const int32_t *__restrict__ probes, int64_t total, int32_t *__restrict__ lcount) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < total) atomicAdd(&lcount[probes[t]], 1);
}

// One workgroup: exclusive scans of the per-cluster pair counts and item counts; writes the
// pair cursor, the items (cluster, first pair, count) and the item total.
__global__ __launch_bounds__(256) void build_items_kernel(const int32_t *__restrict__ lcount, int nl,
                                                          int32_t *__restrict__ loffq, int32_t *__restrict__ cursor,
                                                          int4 *__restrict__ items, int32_t *__restrict__ ntot) {
    __shared__ int32_t s_pairs[256], s_items[256];
    const int tid = threadIdx.x;
    const int per = (nl + 255) / 256;
    const int c0 = tid * per, c1 = min(nl, c0 + per);
    int pairs = 0, its = 0;
    for (int c = c0; c < c1; ++c) {
        pairs += lcount[c];
        its += (lcount[c] + 31) / 32;
    }
    s_pairs[tid] = pairs;
    s_items[tid] = its;
    __syncthreads();
    if (tid == 0) {
        int a = 0, b = 0;
        for (int i = 0; i < 256; ++i) {
            const int pa = s_pairs[i], pb = s_items[i];
            s_pairs[i] = a;
            s_items[i] = b;
            a += pa;
            b += pb;
        }
        *ntot = b;
    }
    __syncthreads();
    int pa = s_pairs[tid], ib = s_items[tid];
    for (int c = c0; c < c1; ++c) {
        const int m = lcount[c];
        loffq[c] = pa;
        cursor[c] = pa;
        for (int s = 0; s < m; s += 32) items[ib++] = make_int4(c, pa + s, min(32, m - s), 0);
        pa += m;
    }
}

__global__ void probe_scatter_kernel(const int32_t *__restrict__ probes, int64_t total, int np,
                                     int32_t *__restrict__ cursor, int32_t *__restrict__ pair_q,
                                     int32_t *__restrict__ pair_p) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < total) {
        const int pos = atomicAdd(&cursor[probes[t]], 1);
        pair_q[pos] = (int32_t)(t / np);
        pair_p[pos] = (int32_t)(t % np);
    }
}

template <int PREC>
__global__ __launch_bounds__(256) void ivfl2_scan_kernel(const __bf16 *__restrict__ Lh, const __bf16 *__restrict__ Ll,
                                                         const float *__restrict__ lxn, const int64_t *__restrict__ loff,
                                                         int dp, const __bf16 *__restrict__ Qh,
                                                         const __bf16 *__restrict__ Ql, const float *__restrict__ qn,
                                                         const int4 *__restrict__ items, const int32_t *__restrict__ ntot,
                                                         const int32_t *__restrict__ pair_q,
                                                         const int32_t *__restrict__ pair_p, int np,
                                                         uint64_t *__restrict__ cand) {
    constexpr int KL = kIvfKL, KP = kIvfKP;
    __shared__ uint64_t lds[8][32][KL];
    __shared__ int ptr[8][32];
    if ((int)blockIdx.x >= *ntot) return; // grid is an upper bound on the item count
    const int4 it = items[blockIdx.x];
    const int c = it.x, start = it.y, cnt = it.z;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, half = lane >> 5, col = lane & 31;
    const int pj = start + min(col, cnt - 1);
    const int64_t q = pair_q[pj];
    const int p = pair_p[pj];
    const int64_t r0 = loff[c], r1 = loff[c + 1];
    const float qnv = qn[q];
    const __bf16 *qh = Qh + q * dp + 8 * half;
    const __bf16 *ql = Ql + q * dp + 8 * half;

    uint32_t Lk[KL], Li[KL];
#pragma unroll
    for (int j = 0; j < KL; ++j) {
        Lk[j] = kKey32None;
        Li[j] = 0u;
    }
    uint32_t thr = kKey32None;

    // Row arrays carry 32 zero rows past the end, so a tile that runs past r1 stays in bounds.
    for (int64_t t = r0 + 32 * w; t < r1; t += 128) {
        float xr[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) xr[r] = lxn[t + 8 * (r >> 2) + 4 * half + (r & 3)];
        const __bf16 *xh = Lh + (t + col) * dp + 8 * half;
        const __bf16 *xl = Ll + (t + col) * dp + 8 * half;
        f32x16 acc = {};
        bf16x8 ah = *(const bf16x8 *)xh, bh = *(const bf16x8 *)qh, al = {}, bl = {};
        if (PREC == 3) {
            al = *(const bf16x8 *)xl;
            bl = *(const bf16x8 *)ql;
        }
        for (int kk = 0; kk < dp; kk += 16) {
            bf16x8 nah = ah, nbh = bh, nal = al, nbl = bl;
            if (kk + 16 < dp) {
                nah = *(const bf16x8 *)(xh + kk + 16);
                nbh = *(const bf16x8 *)(qh + kk + 16);
                if (PREC == 3) {
                    nal = *(const bf16x8 *)(xl + kk + 16);
                    nbl = *(const bf16x8 *)(ql + kk + 16);
                }
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
            if (PREC == 3) {
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
            }
            ah = nah;
            bh = nbh;
            al = nal;
            bl = nbl;
        }
        uint32_t key[16];
        uint32_t mn = kKey32None;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int64_t i = t + 8 * (r >> 2) + 4 * half + (r & 3);
            const float v = fmaxf(0.f, (qnv + xr[r]) - 2.f * acc[r]);
            key[r] = i < r1 ? f32_key(v) : kKey32None;
            mn = min(mn, key[r]);
        }
        while (__any(mn < thr)) {
            if (mn < thr) {
                uint32_t ir = 0;
                bool found = false;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const bool hit = !found && key[r] == mn;
                    ir = hit ? (uint32_t)(t + 8 * (r >> 2) + 4 * half + (r & 3)) : ir;
                    key[r] = hit ? kKey32None : key[r];
                    found = found || hit;
                }
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
                thr = Lk[KL - 1];
            }
            mn = kKey32None;
#pragma unroll
            for (int r = 0; r < 16; ++r) mn = min(mn, key[r]);
        }
    }
#pragma unroll
    for (int j = 0; j < KL; ++j)
        lds[w * 2 + half][col][j] = Lk[j] == kKey32None ? kKey64None : (((uint64_t)Lk[j] << 32) | Li[j]);
    if (threadIdx.x < 256 && lane < 32) ptr[w * 2][col] = 0, ptr[w * 2 + 1][col] = 0;
    __syncthreads();
    if (threadIdx.x < 32 && col < cnt) { // 8-way merge of this query's lists -> its KP best for this probe
        uint64_t *out = cand + ((int64_t)q * np + p) * KP;
        for (int e = 0; e < KP; ++e) {
            int best = 0;
            uint64_t bv = kKey64None;
#pragma unroll
            for (int l = 0; l < 8; ++l) {
                const int pl = ptr[l][col];
                const uint64_t v = pl < KL ? lds[l][col][pl] : kKey64None;
                if (v < bv) {
                    bv = v;
                    best = l;
                }
            }
            if (bv != kKey64None) ptr[best][col] += 1;
            out[e] = bv;
        }
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

// One wave per query: the K2 best candidates by the approximation, exact squared L2 for them,
// the k smallest (distance, id) out.
template <int PER>
__global__ __launch_bounds__(256) void ivfl2_rerank_kernel(const float *__restrict__ X, int64_t d,
                                                           const int32_t *__restrict__ lids, const float *__restrict__ Q,
                                                           int64_t qstride, const uint64_t *__restrict__ cand, int ncand,
                                                           int64_t nq, int k, int32_t *__restrict__ out_idx,
                                                           float *__restrict__ out_dist, int32_t *__restrict__ out_cnt) {
    const int lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= nq) return;
    uint64_t L[PER]; // this lane's candidates, sorted
#pragma unroll
    for (int j = 0; j < PER; ++j) L[j] = kKey64None;
    const uint64_t *src = cand + q * ncand;
    for (int e = lane; e < ncand; e += 64) {
        const uint64_t v = src[e];
        if (v < L[PER - 1]) list_insert_u64<PER>(L, v);
    }
    uint64_t mine = kKey64None; // after the loop lane r holds the r-th best candidate
    for (int r = 0; r < kIvfK2; ++r) {
        const uint64_t m = wmin64(L[0]);
        if (m == kKey64None) break;
        if (L[0] == m) {
#pragma unroll
            for (int j = 0; j < PER - 1; ++j) L[j] = L[j + 1];
            L[PER - 1] = kKey64None;
        }
        if (lane == r) mine = m;
    }
    uint64_t key = kKey64None;
    if (mine != kKey64None) {
        const int32_t id = lids[(uint32_t)(mine & 0xFFFFFFFFull)];
        const float core = l2_core(Q + q * qstride, X + (int64_t)id * d, d);
        key = ((uint64_t)f32_key(core) << 32) | (uint32_t)id;
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
    float *X = (float *)h->X.ensure((size_t)n * d * 4);
    PN_HIP(hipMemcpyAsync(X, dX, (size_t)n * d * 4, hipMemcpyDeviceToDevice, st));
    int32_t *dl = (int32_t *)h->lids.ensure((size_t)n * 4);
    int64_t *dof = (int64_t *)h->loff.ensure((size_t)(nl + 1) * 8);
    PN_HIP(hipMemcpyAsync(dl, ids.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
    PN_HIP(hipMemcpyAsync(dof, off.data(), (size_t)(nl + 1) * 8, hipMemcpyHostToDevice, st));
    // rows in cluster order -> bf16 split (+ zero rows past the end, see the scan kernel)
    DevBuf ordered;
    float *Xo = (float *)ordered.ensure((size_t)n * d * 4);
    hipLaunchKernelGGL(gather_rows_f32_kernel, dim3((unsigned)ceil_div(n * d, 256)), dim3(256), 0, st, X, d, dl, n, Xo);
    l2_prepare(h->lrows, Xo, n, d, d, st);
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
    PN_HIP(hipStreamSynchronize(st)); // host vectors and temporaries are released on return
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
    // bf16 split of the queries
    const int dp = h->lrows.dp;
    __bf16 *Qh = (__bf16 *)h->qsplit_h.ensure((size_t)nq * dp * 2);
    __bf16 *Ql = (__bf16 *)h->qsplit_l.ensure((size_t)nq * dp * 2);
    float *qn = (float *)h->qn.ensure((size_t)nq * 4);
    const int64_t total = nq * np;
    int32_t *lc = (int32_t *)h->lcount.ensure((size_t)h->nl * 4);
    int32_t *lo = (int32_t *)h->loffq.ensure((size_t)h->nl * 4);
    int32_t *cur = (int32_t *)h->cursor.ensure((size_t)h->nl * 4);
    int32_t *pq = (int32_t *)h->pair_q.ensure((size_t)total * 4);
    int32_t *pp = (int32_t *)h->pair_p.ensure((size_t)total * 4);
    const int64_t max_items = ceil_div(total, 32) + h->nl;
    int4 *items = (int4 *)h->items.ensure((size_t)max_items * 16);
    int32_t *ntot = (int32_t *)h->ntot.ensure(4);
    uint64_t *cand = (uint64_t *)h->cand.ensure((size_t)total * kIvfKP * 8);
    {
        StageScope sc(st, STAGE_OTHER);
        l2_split_rows(Q, nq, d, re, dp, Qh, Ql, qn, st);
        PN_HIP(hipMemsetAsync(lc, 0, (size_t)h->nl * 4, st));
        hipLaunchKernelGGL(probe_count_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, st, pr, total, lc);
        hipLaunchKernelGGL(build_items_kernel, dim3(1), dim3(256), 0, st, lc, h->nl, lo, cur, items, ntot);
        hipLaunchKernelGGL(probe_scatter_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, st, pr, total, np,
                           cur, pq, pp);
    }
    {
        StageScope sc(st, STAGE_SCAN);
        if (prec == 3)
            hipLaunchKernelGGL((ivfl2_scan_kernel<3>), dim3((unsigned)max_items), dim3(256), 0, st,
                               h->lrows.hi.as<__bf16>(), h->lrows.lo.as<__bf16>(), h->lrows.norm.as<float>(),
                               h->loff.as<int64_t>(), dp, Qh, Ql, qn, items, ntot, pq, pp, np, cand);
        else
            hipLaunchKernelGGL((ivfl2_scan_kernel<1>), dim3((unsigned)max_items), dim3(256), 0, st,
                               h->lrows.hi.as<__bf16>(), h->lrows.lo.as<__bf16>(), h->lrows.norm.as<float>(),
                               h->loff.as<int64_t>(), dp, Qh, Ql, qn, items, ntot, pq, pp, np, cand);
        PN_HIP(hipGetLastError());
    }
    {
        StageScope sc(st, STAGE_RERANK);
        const int ncand = np * kIvfKP;
        dim3 g((unsigned)ceil_div(nq, 4));
        if (ncand <= 64 * 8)
            hipLaunchKernelGGL((ivfl2_rerank_kernel<8>), g, dim3(256), 0, st, h->X.as<float>(), d, h->lids.as<int32_t>(),
                               Q, re, cand, ncand, nq, k, oi, od, oc);
        else if (ncand <= 64 * 16)
            hipLaunchKernelGGL((ivfl2_rerank_kernel<16>), g, dim3(256), 0, st, h->X.as<float>(), d,
                               h->lids.as<int32_t>(), Q, re, cand, ncand, nq, k, oi, od, oc);
        else
            hipLaunchKernelGGL((ivfl2_rerank_kernel<32>), g, dim3(256), 0, st, h->X.as<float>(), d,
                               h->lids.as<int32_t>(), Q, re, cand, ncand, nq, k, oi, od, oc);
        PN_HIP(hipGetLastError());
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
        PN_REQUIRE(k <= 64, PN_EINVAL, "IVFFlatL2Index search supports k <= 64");
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
        ivfl2_search_enqueue(h, q, d, nq, nprobe, k, precision == 1 ? 3 : 1, i32, f32, c32, st);
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
        ivfl2_search_enqueue(h, queries, h->d, nq, nprobe, k, precision == 1 ? 3 : 1, out_idx, out_dist, out_count,
                             (hipStream_t)stream);
    });
}
