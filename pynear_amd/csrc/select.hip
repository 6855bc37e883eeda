// Selection stages of the exact search.
//
//  * merge_lists: the per-slice top-K lists of a scan are merged into kk results per query,
//    ascending (distance, index) -- the total order that replaces the reference's max-heaps
//    (VPTree.hpp:419-424, :469-473) whose tie order is unspecified. One lane per query; each
//    slice's list is sorted, so a lane stops reading a slice at its first key that cannot enter.
//  * matrix_search: any k and any row width -- a distance matrix for a chunk of queries and a
//    stable segmented radix sort of (distance key, index) per query; kk columns are kept.
//  * pairwise_distance: row-wise distances for the free functions dist_l2/.../dist_hamming_*.
#include <hipcub/hipcub.hpp>

#include "common.hpp"
#include "device_util.hpp"
#include "kernels.hpp"
#include "merge_dev.hpp"
#include "ref_arith.hpp"

namespace pn {

template <int K>
__global__ __launch_bounds__(256) void merge_lists_kernel(const uint64_t *__restrict__ part, int64_t slices,
                                                          int64_t nq, int64_t nq_pad, int kk, int kind,
                                                          int32_t *__restrict__ out_idx, void *__restrict__ out_dist) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    uint64_t L[K];
#pragma unroll
    for (int j = 0; j < K; ++j) L[j] = kKey64None;
    for (int64_t s = 0; s < slices; ++s) {
        const uint64_t *src = part + (s * K) * nq_pad + q;
        for (int j = 0; j < K; ++j) {
            uint64_t key = src[(int64_t)j * nq_pad];
            if (key == kKey64None) break;
            if (kind == OUT_L2_F32) {
                // order by the final (sqrt) distance; sqrt is monotone, so the slice stays sorted
                // in the high word, but keys whose roots tie may follow in any index order: the
                // slice is left only past the last key's high word
                float dd = sqrtf(key_f32((uint32_t)(key >> 32)));
                key = ((uint64_t)f32_key(dd) << 32) | (key & 0xFFFFFFFFull);
                if (key >= L[K - 1]) {
                    if ((key >> 32) > (L[K - 1] >> 32)) break;
                    continue;
                }
            }
            if (key >= L[K - 1]) break;
            list_insert_u64<K>(L, key);
        }
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (j < kk) {
            out_idx[q * kk + j] = (int32_t)(uint32_t)(L[j] & 0xFFFFFFFFull);
            uint32_t hi = (uint32_t)(L[j] >> 32);
            if (kind == OUT_HAMMING_I32)
                ((int32_t *)out_dist)[q * kk + j] = (int32_t)hi;
            else
                ((float *)out_dist)[q * kk + j] = key_f32(hi);
        }
    }
}

template <int K>
__global__ __launch_bounds__(256) void merge_lists_wave_kernel(const uint64_t *__restrict__ part, int64_t slices,
                                                               int klist, int64_t nq, int64_t qs, int64_t es, int kk,
                                                               int kind, int32_t *__restrict__ out_idx,
                                                               void *__restrict__ out_dist) {
    __shared__ uint64_t wbuf[4][256];
    const int64_t q = xcd_contiguous_block() * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (q >= nq) return; // whole wave exits together
    uint64_t unused;
    merge_wave_query<K>(part, slices, klist, qs, es, q, kk, kind, out_idx, out_dist, wbuf[threadIdx.x >> 6], unused);
}

__global__ __launch_bounds__(256) void merge_lists_lean_kernel(const uint64_t *__restrict__ part, int64_t slices,
                                                               int klist, int64_t nq, int64_t qs, int64_t es, int kk,
                                                               int kind, int32_t *__restrict__ out_idx,
                                                               void *__restrict__ out_dist, bool rounds_only) {
    __shared__ uint64_t wbuf[4][256];
    const int64_t q = xcd_contiguous_block() * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (q >= nq) return; // whole wave exits together
    uint64_t unused;
    merge_wave_lean(part, slices, klist, qs, es, q, kk, kind, out_idx, out_dist, wbuf[threadIdx.x >> 6], rounds_only,
                    unused);
}

// merge.lean: 1 for private-list capacities of 16 or more (below that the private lists cost
// little and the lean merge's extra pass shows), 3 always, 0 never, 2 always with the rounds path
// only (a test setting)
static bool lean_merge(int kk, int cap) {
    const int64_t t = tune_get("merge.lean");
    return kk <= 64 && (t >= 2 || (t == 1 && cap >= 16));
}
static bool lean_rounds() { return tune_get("merge.lean") == 2; }

void merge_lists(const uint64_t *part, int64_t slices, int K, int64_t nq, int64_t nq_pad, int kk, int kind,
                 int32_t *out_idx, void *out_dist, hipStream_t st) {
    const bool wave = tune_get("merge.wave_per_query") != 0;
    if (wave && lean_merge(kk, K)) {
        hipLaunchKernelGGL(merge_lists_lean_kernel, dim3(xcd_grid(ceil_div(nq, 4))), dim3(256), 0, st, part, slices,
                           K, nq, (int64_t)1, nq_pad, kk, kind, out_idx, out_dist, lean_rounds());
        PN_HIP(hipGetLastError());
        return;
    }
    dim3 grid(wave ? xcd_grid(ceil_div(nq, 4)) : (unsigned)ceil_div(nq, 256));
#define PN_MERGE(KV)                                                                                   \
    do {                                                                                               \
        if (wave)                                                                                      \
            hipLaunchKernelGGL((merge_lists_wave_kernel<KV>), grid, dim3(256), 0, st, part, slices, KV, nq,    \
                               (int64_t)1, nq_pad, kk, kind, out_idx, out_dist);                       \
        else                                                                                           \
            hipLaunchKernelGGL((merge_lists_kernel<KV>), grid, dim3(256), 0, st, part, slices, nq, nq_pad, kk, \
                               kind, out_idx, out_dist);                                               \
    } while (0)
    switch (K) {
    case 1: PN_MERGE(1); break;
    case 2: PN_MERGE(2); break;
    case 4: PN_MERGE(4); break;
    case 8: PN_MERGE(8); break;
    case 16: PN_MERGE(16); break;
    case 32: PN_MERGE(32); break;
    default: fail(PN_EINVAL, "merge: unsupported list size");
    }
#undef PN_MERGE
    PN_HIP(hipGetLastError());
}

void merge_lists_wave(const uint64_t *part, int64_t slices, int klist, int cap, int64_t nq, int64_t qs, int64_t es,
                      int kk, int kind, int32_t *out_idx, void *out_dist, hipStream_t st) {
    dim3 grid(xcd_grid(ceil_div(nq, 4)));
    if (lean_merge(kk, cap)) {
        hipLaunchKernelGGL(merge_lists_lean_kernel, grid, dim3(256), 0, st, part, slices, klist, nq, qs, es, kk, kind,
                           out_idx, out_dist, lean_rounds());
        PN_HIP(hipGetLastError());
        return;
    }
#define PN_MERGEW(KV)                                                                                        \
    hipLaunchKernelGGL((merge_lists_wave_kernel<KV>), grid, dim3(256), 0, st, part, slices, klist, nq, qs, es, kk, \
                       kind, out_idx, out_dist)
    switch (cap) {
    case 1: PN_MERGEW(1); break;
    case 2: PN_MERGEW(2); break;
    case 4: PN_MERGEW(4); break;
    case 8: PN_MERGEW(8); break;
    case 16: PN_MERGEW(16); break;
    case 32: PN_MERGEW(32); break;
    case 64: PN_MERGEW(64); break;
    default: fail(PN_EINVAL, "merge: unsupported list capacity");
    }
#undef PN_MERGEW
    PN_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- distance matrix path
// D[qi, i] = monotone u32 key of the final distance (bits of the float, or the Hamming count).
// With a candidate filter (group words gw > 0) a Hamming row that is no candidate gets the empty
// key, which sorts last and comes back as index -1.
__global__ void dist_matrix_kernel(int metric, const void *__restrict__ db, int64_t n, int64_t row_elems,
                                   int64_t dim, const void *__restrict__ qs, int64_t q0, int64_t nqc,
                                   uint32_t *__restrict__ D, int32_t *__restrict__ V, int gw, int subs, int rsub) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t qi = blockIdx.y;
    if (i >= n || qi >= nqc) return;
    uint32_t key;
    if (metric == PN_METRIC_HAMMING) {
        const uint32_t *a = (const uint32_t *)qs + (q0 + qi) * row_elems;
        const uint32_t *b = (const uint32_t *)db + i * row_elems;
        uint32_t acc = 0;
        if (gw > 0) {
            int mn = INT32_MAX;
            for (int64_t g = 0; g * gw < row_elems; ++g) {
                uint32_t s = 0;
                for (int j = 0; j < gw; ++j) s += __popc(a[g * gw + j] ^ b[g * gw + j]);
                acc += s;
                if (g < subs) mn = min(mn, (int)s);
            }
            key = mn <= rsub ? acc : kKey32None;
        } else {
            for (int64_t w = 0; w < row_elems; ++w) acc = __popc(a[w] ^ b[w]) + acc;
            key = acc;
        }
    } else {
        const float *a = (const float *)qs + (q0 + qi) * row_elems;
        const float *b = (const float *)db + i * row_elems;
        key = f32_key(ref_distance(metric, a, b, dim));
    }
    D[qi * n + i] = key;
    V[qi * n + i] = (int32_t)i;
}

__global__ void fill_offsets_kernel(int64_t *off, int64_t nseg, int64_t n) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s <= nseg) off[s] = s * n;
}

__global__ void gather_topk_kernel(const uint32_t *__restrict__ Ds, const int32_t *__restrict__ Vs, int64_t n,
                                   int64_t nqc, int64_t q0, int kk, int metric, int32_t *__restrict__ out_idx,
                                   void *__restrict__ out_dist) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nqc * kk) return;
    const int64_t qi = t / kk, j = t % kk;
    const int64_t o = (q0 + qi) * kk + j;
    uint32_t key = Ds[qi * n + j];
    out_idx[o] = metric == PN_METRIC_HAMMING && key == kKey32None ? -1 : Vs[qi * n + j];
    if (metric == PN_METRIC_HAMMING)
        ((int32_t *)out_dist)[o] = (int32_t)key;
    else
        ((float *)out_dist)[o] = key_f32(key);
}

static size_t sort_temp_bytes(int64_t n, int64_t q_chunk) {
    size_t tmp = 0;
    int64_t items = n * q_chunk;
    (void)hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, tmp, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                      (const int32_t *)nullptr, (int32_t *)nullptr, (int)items,
                                                      (int)q_chunk, (const int64_t *)nullptr,
                                                      (const int64_t *)nullptr);
    return tmp;
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

size_t matrix_search_workspace(int64_t n, int64_t q_chunk) {
    size_t items = (size_t)n * q_chunk;
    return 4 * align256(items * 4) + align256((q_chunk + 1) * 8) + align256(sort_temp_bytes(n, q_chunk));
}

void matrix_search(int metric, const void *db, int64_t n, int64_t row_elems, int64_t dim, const void *queries,
                   int64_t nq, int kk, int32_t *out_idx, void *out_dist, void *ws, size_t ws_bytes, int64_t q_chunk,
                   hipStream_t st, const SubFilter &filt) {
    PN_REQUIRE(n * q_chunk < (int64_t)INT32_MAX, PN_EINVAL, "matrix search: chunk too large");
    PN_REQUIRE(filt.group_words == 0 || (metric == PN_METRIC_HAMMING && row_elems % filt.group_words == 0), PN_EINVAL,
               "matrix search: the candidate filter needs Hamming rows of whole groups");
    size_t items = (size_t)n * q_chunk;
    char *p = (char *)ws;
    uint32_t *D = (uint32_t *)p;          p += align256(items * 4);
    uint32_t *Ds = (uint32_t *)p;         p += align256(items * 4);
    int32_t *V = (int32_t *)p;            p += align256(items * 4);
    int32_t *Vs = (int32_t *)p;           p += align256(items * 4);
    int64_t *off = (int64_t *)p;          p += align256((q_chunk + 1) * 8);
    size_t tmp_bytes = sort_temp_bytes(n, q_chunk);
    void *tmp = p;
    PN_REQUIRE((size_t)(p - (char *)ws) + tmp_bytes <= ws_bytes, PN_ERUNTIME, "matrix search: workspace too small");
    for (int64_t q0 = 0; q0 < nq; q0 += q_chunk) {
        int64_t nqc = std::min(q_chunk, nq - q0);
        dim3 g1((unsigned)ceil_div(n, 256), (unsigned)nqc);
        hipLaunchKernelGGL(dist_matrix_kernel, g1, dim3(256), 0, st, metric, db, n, row_elems, dim, queries, q0, nqc,
                           D, V, filt.group_words, filt.subs, filt.rsub);
        hipLaunchKernelGGL(fill_offsets_kernel, dim3((unsigned)ceil_div(nqc + 1, 256)), dim3(256), 0, st, off, nqc, n);
        size_t tb = tmp_bytes;
        PN_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(tmp, tb, D, Ds, V, Vs, (int)(n * nqc), (int)nqc, off,
                                                           off + 1, 0, 32, st));
        hipLaunchKernelGGL(gather_topk_kernel, dim3((unsigned)ceil_div(nqc * kk, 256)), dim3(256), 0, st, Ds, Vs, n,
                           nqc, q0, kk, metric, out_idx, out_dist);
        PN_HIP(hipGetLastError());
    }
}

// ---------------------------------------------------------------- pairwise distances
__global__ void pairwise_kernel(int metric, const void *__restrict__ a, const void *__restrict__ b, int64_t n,
                                int64_t dim, void *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (metric == PN_METRIC_HAMMING) {
        const uint8_t *x = (const uint8_t *)a + i * dim, *y = (const uint8_t *)b + i * dim;
        int64_t h = 0;
        for (int64_t e = 0; e < dim; ++e) h += __popc((uint32_t)(x[e] ^ y[e]));
        ((int64_t *)out)[i] = h;
    } else {
        ((float *)out)[i] = ref_distance(metric, (const float *)a + i * dim, (const float *)b + i * dim, dim);
    }
}

void pairwise_distance(int metric, const void *a, const void *b, int64_t n, int64_t dim, void *out, hipStream_t st) {
    hipLaunchKernelGGL(pairwise_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, metric, a, b, n, dim, out);
    PN_HIP(hipGetLastError());
}

} // namespace pn
