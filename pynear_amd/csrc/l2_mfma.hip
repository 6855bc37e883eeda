// Exact L2 k-NN through a bf16 MFMA candidate scan and an exact rerank.
//
// The reference computes every distance as sqrt(sum (q-x)^2) (DistanceFunctions.hpp:304-345). A
// full scan in that form costs two VALU operations per element; here the scan is the GEMM form
// |q|^2 + |x|^2 - 2 q.x on the matrix cores, and only a few candidates per query are recomputed
// in the reference's arithmetic:
//   1. split: x = hi + lo with hi = bf16(x), lo = bf16(x - hi); |x|^2 accumulated in fp64.
//   2. scan: v_mfma_f32_32x32x16_bf16 over (32 rows x 32 queries) tiles, three products per
//      16-wide K step (hi.hi + hi.lo + lo.hi; relative error of the dot product ~2^-16), fp32
//      accumulate. Each lane owns one query and a sorted list of its KL best (approx, row)
//      candidates; lanes of every (wave, half-wave) form one partial list.
//   3. merge: the partial lists give each query its K2 best candidates by the approximation.
//   4. rerank: the K2 candidates are recomputed exactly (reference order) and ordered by
//      (distance, index). A candidate set is proven complete when the smallest approximation any
//      list or the merge could have dropped, minus a rigorous error bound, exceeds the k-th exact
//      squared distance; otherwise the query is flagged.
//   5. fallback: flagged queries (rare) are searched by an exact full scan on the device.
//
// MFMA layout (checked on the device): A 32x16 -- lane l holds row l%32, k = 8(l/32)..+7;
// B 16x32 -- lane l holds column l%32, k = 8(l/32)..+7; D 32x32 -- lane l, register r holds
// row 8(r/4) + 4(l/32) + r%4, column l%32.
#include <type_traits>

#include "common.hpp"
#include "device_util.hpp"
#include "kernels.hpp"
#include "l2_mfma.hpp"
#include "ref_arith.hpp"

namespace pn {


// Rows [n, n_out) are written as zero rows (padding that later kernels read unguarded).
__global__ __launch_bounds__(256) void l2_split_kernel(const float *__restrict__ X, int64_t n, int64_t n_out,
                                                       int64_t d, int64_t stride, int dp, __bf16 *__restrict__ H,
                                                       __bf16 *__restrict__ Lo, float *__restrict__ nrm) {
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n_out) return;
    double acc = 0.0;
    for (int e = lane; e < dp; e += 64) {
        const float x = e < d && row < n ? X[row * stride + e] : 0.f;
        const __bf16 h = (__bf16)x;
        const __bf16 l = (__bf16)(x - (float)h);
        H[row * dp + e] = h;
        Lo[row * dp + e] = l;
        acc += (double)x * (double)x;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) nrm[row] = (float)acc;
}

__global__ void max_f32_kernel(const float *__restrict__ v, int64_t n, unsigned int *__restrict__ out) {
    __shared__ float red[256];
    float m = 0.f;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) m = fmaxf(m, v[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, __float_as_uint(red[0])); // non-negative floats order as uints
}

// Keys of one 32 x 32 tile for this lane's query (rows t + 8g + 4 half + {0..3}, g < 4) into
// the lane's sorted list (Lk, Li) of KL. A key is the approximation's float bits with the low 5
// bits replaced by the row's offset in the tile: keys stay unique inside a tile, the minimum names
// its row, and (key & ~31) is a lower bound of the approximation. Passing keys (below thr) are
// inserted one per lane per round; rounds run while any lane of the wave still has one. thr may
// come from the other half-wave's list of the same query: rows at or above it cannot reach the
// KL best of the two lists together.
template <int KL, bool MASK, bool FIRST>
__device__ __forceinline__ void tile_keys_insert(const f32x16 &acc, const float4 (&xv)[4], float qnv, int64_t t,
                                                 int64_t r1, int half, uint32_t (&Lk)[KL], uint32_t (&Li)[KL],
                                                 uint32_t &thr) {
    uint32_t key[16];
    uint32_t mn = kKey32None;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int off = 8 * (r >> 2) + 4 * half + (r & 3);
        const float xr = (r & 3) == 0 ? xv[r >> 2].x : (r & 3) == 1 ? xv[r >> 2].y : (r & 3) == 2 ? xv[r >> 2].z : xv[r >> 2].w;
        const float v = fmaxf(0.f, (qnv + xr) - 2.f * acc[r]);
        key[r] = !MASK || t + off < r1 ? ((f32_key(v) & ~31u) | (uint32_t)off) : kKey32None;
        mn = min(mn, key[r]);
    }
    if (FIRST) { // empty list: the tile's KL smallest keys, sorted, by a network instead of KL rounds
        sort16_u32(key);
#pragma unroll
        for (int j = 0; j < KL; ++j) {
            Lk[j] = j < 16 ? key[j] : kKey32None;
            Li[j] = j < 16 ? (uint32_t)t + (key[j] & 31u) : 0u;
        }
        thr = min(thr, Lk[KL - 1]);
        return;
    }
    while (__any(mn < thr)) {
        if (mn < thr) {
#pragma unroll
            for (int r = 0; r < 16; ++r) key[r] = key[r] == mn ? kKey32None : key[r];
            const uint32_t ir = (uint32_t)t + (mn & 31u);
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

// Half-wave merge: lanes l and l+32 hold lists of the same query; lane l writes the KL smallest
// of both to out[j * nq_pad]. Keys a list dropped are >= its last key >= the merged last key, so
// the merged last key still bounds every dropped candidate from below.
template <int KL>
__device__ __forceinline__ void half_merge_out(uint64_t (*lds)[KL], int half, const uint32_t (&Lk)[KL],
                                               const uint32_t (&Li)[KL], uint64_t *out, int64_t nq_pad) {
    if (half == 1) {
#pragma unroll
        for (int j = 0; j < KL; ++j) lds[0][j] = Lk[j] == kKey32None ? kKey64None : (((uint64_t)Lk[j] << 32) | Li[j]);
    }
    __syncthreads();
    if (half == 0) {
        int pb = 0;
        int pa = 0;
        for (int j = 0; j < KL; ++j) {
            uint64_t va = kKey64None;
#pragma unroll
            for (int e = 0; e < KL; ++e) // own list entry pa, selected without dynamic register indexing
                if (e == pa) va = Lk[e] == kKey32None ? kKey64None : (((uint64_t)Lk[e] << 32) | Li[e]);
            const uint64_t vb = pb < KL ? lds[0][pb] : kKey64None;
            const bool takea = va <= vb;
            out[(int64_t)j * nq_pad] = takea ? va : vb;
            pa += takea ? 1 : 0;
            pb += takea ? 0 : 1;
        }
    }
}

template <int KL, int PREC>
__global__ __launch_bounds__(256) void l2_mfma_scan_kernel(const __bf16 *__restrict__ Xh, const __bf16 *__restrict__ Xl,
                                                           const float *__restrict__ xn, int64_t n, int dp,
                                                           const __bf16 *__restrict__ Qh, const __bf16 *__restrict__ Ql,
                                                           const float *__restrict__ qn, int64_t rows_per_slice,
                                                           uint64_t *__restrict__ part, int64_t nq_pad) {
    // A workgroup owns 128 queries (wave w: queries 32w..32w+31 of the block) and one slice of
    // rows; the four waves walk the same 32-row tiles, so each tile is fetched from L2 once per
    // workgroup and served to the other waves from L1. Row arrays are padded with zero rows to a
    // multiple of 32 and slices start on 32-row boundaries: tile loads need no guard, rows >= r1
    // are masked.
    __shared__ uint64_t lds[4][32][KL];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, half = lane >> 5, col = lane & 31;
    const int64_t q = (int64_t)blockIdx.x * 128 + 32 * w + col; // query buffers are padded to nq_pad rows
    const int64_t s = blockIdx.y;
    const int64_t r0 = s * rows_per_slice, r1 = min(n, r0 + rows_per_slice);
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

    for (int64_t t = r0; t < r1; t += 32) {
        // |x|^2 of this lane's 16 output rows: rows t + 8g + 4 half + {0..3}
        float4 xv[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) xv[g] = *(const float4 *)(xn + t + 8 * g + 4 * half);
        const __bf16 *xh = Xh + (t + col) * dp + 8 * half;
        const __bf16 *xl = Xl + (t + col) * dp + 8 * half;
        f32x16 acc = {};
        mfma_tile_bf16<PREC, PREC == 3 ? 4 : 8>(acc, xh, xl, qh, ql, dp);
        thr = min(thr, (uint32_t)__shfl_xor((int)thr, 32, 64)); // the other half-wave's list of this query
        if (t == r0) // only a slice's last tile can hold rows past its end
            t + 32 <= r1 ? tile_keys_insert<KL, false, true>(acc, xv, qnv, t, r1, half, Lk, Li, thr)
                         : tile_keys_insert<KL, true, true>(acc, xv, qnv, t, r1, half, Lk, Li, thr);
        else
            t + 32 <= r1 ? tile_keys_insert<KL, false, false>(acc, xv, qnv, t, r1, half, Lk, Li, thr)
                         : tile_keys_insert<KL, true, false>(acc, xv, qnv, t, r1, half, Lk, Li, thr);
    }
    half_merge_out<KL>(lds[w] + col, half, Lk, Li, part + (s * KL) * nq_pad + q, nq_pad);
}

// The same scan for dp = 16 NS (NS <= 4): the lane's query fragments stay in registers for the
// whole slice, and the next tile's row fragments and norms are loaded before the current tile's
// products and keys, so the memory latency of one tile hides behind the work of the previous one.
template <int KL, int PREC, int NS>
__global__ __launch_bounds__(256) void l2_mfma_scan_small_kernel(
    const __bf16 *__restrict__ Xh, const __bf16 *__restrict__ Xl, const float *__restrict__ xn, int64_t n,
    const __bf16 *__restrict__ Qh, const __bf16 *__restrict__ Ql, const float *__restrict__ qn,
    int64_t rows_per_slice, uint64_t *__restrict__ part, int64_t nq_pad) {
    constexpr int dp = 16 * NS;
    constexpr int NPL = PREC == 3 ? 2 : 1;
    __shared__ uint64_t lds[4][32][KL];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, half = lane >> 5, col = lane & 31;
    const int64_t q = (int64_t)blockIdx.x * 128 + 32 * w + col;
    const int64_t s = blockIdx.y;
    const int64_t r0 = s * rows_per_slice, r1 = min(n, r0 + rows_per_slice);
    const float qnv = qn[q];
    bf16x8 b[NS][NPL];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        b[k][0] = *(const bf16x8 *)(Qh + q * dp + 16 * k + 8 * half);
        if (PREC == 3) b[k][NPL - 1] = *(const bf16x8 *)(Ql + q * dp + 16 * k + 8 * half);
    }
    uint32_t Lk[KL], Li[KL];
#pragma unroll
    for (int j = 0; j < KL; ++j) {
        Lk[j] = kKey32None;
        Li[j] = 0u;
    }
    uint32_t thr = kKey32None;
    bf16x8 a[NS][NPL], an[NS][NPL];
    float4 xv[4], xvn[4];
    auto load_tile = [&](int64_t t, bf16x8 (&fa)[NS][NPL], float4 (&fx)[4]) {
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            fa[k][0] = *(const bf16x8 *)(Xh + (t + col) * dp + 16 * k + 8 * half);
            if (PREC == 3) fa[k][NPL - 1] = *(const bf16x8 *)(Xl + (t + col) * dp + 16 * k + 8 * half);
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) fx[g] = *(const float4 *)(xn + t + 8 * g + 4 * half);
    };
    if (r0 < r1) load_tile(r0, a, xv);
    for (int64_t t = r0; t < r1; t += 32) {
        if (t + 32 < r1) load_tile(t + 32, an, xvn); // padded arrays: in bounds
        f32x16 acc = {};
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[k][0], b[k][0], acc, 0, 0, 0);
            if (PREC == 3) {
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[k][0], b[k][NPL - 1], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[k][NPL - 1], b[k][0], acc, 0, 0, 0);
            }
        }
        thr = min(thr, (uint32_t)__shfl_xor((int)thr, 32, 64)); // the other half-wave's list of this query
        if (t == r0) // only a slice's last tile can hold rows past its end
            t + 32 <= r1 ? tile_keys_insert<KL, false, true>(acc, xv, qnv, t, r1, half, Lk, Li, thr)
                         : tile_keys_insert<KL, true, true>(acc, xv, qnv, t, r1, half, Lk, Li, thr);
        else
            t + 32 <= r1 ? tile_keys_insert<KL, false, false>(acc, xv, qnv, t, r1, half, Lk, Li, thr)
                         : tile_keys_insert<KL, true, false>(acc, xv, qnv, t, r1, half, Lk, Li, thr);
#pragma unroll
        for (int k = 0; k < NS; ++k)
#pragma unroll
            for (int pl = 0; pl < NPL; ++pl) a[k][pl] = an[k][pl];
#pragma unroll
        for (int g = 0; g < 4; ++g) xv[g] = xvn[g];
    }
    half_merge_out<KL>(lds[w] + col, half, Lk, Li, part + (s * KL) * nq_pad + q, nq_pad);
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}

// One wave per query: exact rerank of K2c candidates, (distance, index) order, completeness check.
__global__ __launch_bounds__(256) void l2_rerank_kernel(const float *__restrict__ X, int64_t xstride, int64_t d,
                                                        const float *__restrict__ Q, int64_t qstride,
                                                        const float *__restrict__ qn, const unsigned int *__restrict__ xn_max_bits,
                                                        const int32_t *__restrict__ cidx, const float *__restrict__ capx,
                                                        int K2c, const uint64_t *__restrict__ part, int64_t nlists,
                                                        int KL, int64_t nq_pad, int64_t nq, int kk, int64_t n,
                                                        float eps_coef, int32_t *__restrict__ out_idx,
                                                        float *__restrict__ out_dist, int32_t *__restrict__ flags) {
    const int lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= nq) return;
    // the smallest approximation any full list dropped: its last key with the row bits cleared
    uint32_t bmin = kKey32None;
    for (int64_t l = lane; l < nlists; l += 64) {
        const uint32_t last = (uint32_t)(part[(l * KL + KL - 1) * nq_pad + q] >> 32);
        bmin = min(bmin, last == kKey32None ? kKey32None : (last & ~31u));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) bmin = min(bmin, (uint32_t)__shfl_xor((int)bmin, off, 64));

    uint64_t key = kKey64None;
    float core = 0.f;
    if (lane < K2c) {
        const int32_t idx = cidx[q * K2c + lane];
        if (idx >= 0 && idx < n) {
            core = l2_core(Q + q * qstride, X + (int64_t)idx * xstride, d);
            key = ((uint64_t)f32_key(sqrtf(core)) << 32) | (uint32_t)idx;
        }
    }
    float dk_core = 0.f;
    for (int r = 0; r < kk; ++r) {
        const uint64_t m = wave_min_u64(key);
        const unsigned long long own = __ballot(key == m);
        const int owner = __ffsll((unsigned long long)own) - 1;
        const float cr = __shfl(core, owner, 64);
        if (key == m) key = kKey64None;
        if (lane == 0) {
            out_idx[q * kk + r] = (int32_t)(uint32_t)(m & 0xFFFFFFFFull);
            out_dist[q * kk + r] = key_f32((uint32_t)(m >> 32));
        }
        dk_core = cr;
    }
    if (lane == 0) {
        int flag = 0;
        if (n > K2c) {
            const uint32_t al = __float_as_uint(capx[q * K2c + K2c - 1]);
            const uint32_t bnd = min(bmin, al == kKey32None ? kKey32None : (al & ~31u));
            const float bound = bnd == kKey32None ? INFINITY : key_f32(bnd);
            const float s = sqrtf(qn[q]) + sqrtf(__uint_as_float(*xn_max_bits));
            const float eps = eps_coef * s * s;
            flag = !(bound - eps > dk_core);
        }
        flags[q] = flag;
    }
}

// Exact full scan for flagged queries (one workgroup each; unflagged workgroups exit at once).
template <int K>
__global__ __launch_bounds__(256) void l2_fallback_kernel(const float *__restrict__ X, int64_t xstride, int64_t n,
                                                          int64_t d, const float *__restrict__ Q, int64_t qstride,
                                                          const int32_t *__restrict__ flags, int kk,
                                                          int32_t *__restrict__ out_idx, float *__restrict__ out_dist) {
    __shared__ uint64_t wres[4][K];
    const int64_t q = blockIdx.x;
    if (!flags[q]) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t L[K];
#pragma unroll
    for (int j = 0; j < K; ++j) L[j] = kKey64None;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const float dd = sqrtf(l2_core(Q + q * qstride, X + i * xstride, d));
        const uint64_t key = ((uint64_t)f32_key(dd) << 32) | (uint32_t)i;
        if (key < L[K - 1]) list_insert_u64<K>(L, key);
    }
    for (int r = 0; r < kk; ++r) { // wave-level extraction of the wave's kk best
        const uint64_t m = wave_min_u64(L[0]);
        if (L[0] == m && m != kKey64None) {
#pragma unroll
            for (int j = 0; j < K - 1; ++j) L[j] = L[j + 1];
            L[K - 1] = kKey64None;
        }
        if (lane == 0) wres[w][r] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int p[4] = {0, 0, 0, 0};
        auto head = [&](int ww) { return p[ww] < kk ? wres[ww][p[ww]] : kKey64None; };
        for (int r = 0; r < kk; ++r) {
            int bw = 0;
            for (int ww = 1; ww < 4; ++ww)
                if (head(ww) < head(bw)) bw = ww;
            const uint64_t m = head(bw);
            ++p[bw];
            out_idx[q * kk + r] = (int32_t)(uint32_t)(m & 0xFFFFFFFFull);
            out_dist[q * kk + r] = key_f32((uint32_t)(m >> 32));
        }
    }
}

// ---------------------------------------------------------------- host side
void l2_prepare(L2Set &s, const float *X, int64_t n, int64_t d, int64_t stride, hipStream_t st, int align) {
    s.n = n;
    s.d = d;
    s.dp = (int)(ceil_div(d, (int64_t)align) * align);
    // 32 zero rows past the last full tile: tiles that start at any row < n stay in bounds
    const int64_t npad = ceil_div(n, 32) * 32 + 32;
    __bf16 *H = (__bf16 *)s.hi.ensure((size_t)npad * s.dp * 2);
    __bf16 *L = (__bf16 *)s.lo.ensure((size_t)npad * s.dp * 2);
    float *nr = (float *)s.norm.ensure((size_t)npad * 4);
    unsigned int *mx = (unsigned int *)s.norm_max.ensure(4);
    PN_HIP(hipMemsetAsync(mx, 0, 4, st));
    if (n == 0) return;
    // zero tail rows: tiles read them unguarded
    hipLaunchKernelGGL(l2_split_kernel, dim3((unsigned)ceil_div(npad, 4)), dim3(256), 0, st, X, n, npad, d, stride,
                       s.dp, H, L, nr);
    hipLaunchKernelGGL(max_f32_kernel, dim3((unsigned)std::min<int64_t>(1024, ceil_div(n, 256))), dim3(256), 0, st, nr,
                       n, mx);
    PN_HIP(hipGetLastError());
}

void l2_split_rows(const float *X, int64_t n, int64_t d, int64_t stride, int dp, __bf16 *H, __bf16 *L, float *nrm,
                   hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(l2_split_kernel, dim3((unsigned)ceil_div(n, 4)), dim3(256), 0, st, X, n, n, d, stride, dp, H, L,
                       nrm);
    PN_HIP(hipGetLastError());
}

static int kl_for(int kk) { return kk <= 6 ? 8 : (kk <= 12 ? 16 : 32); }

void l2_mfma_search(const L2Set &xs, const float *X, int64_t xstride, const float *Q, int64_t qstride, int64_t nq,
                    int kk, int32_t *out_idx, float *out_dist, L2Work &wk, int cus, hipStream_t st) {
    const int64_t n = xs.n, d = xs.d;
    const int dp = xs.dp;
    const int KL = kl_for(kk);
    const int K2 = 2 * KL;
    const int K2c = (int)std::min<int64_t>(K2, n);
    const int64_t gx = ceil_div(nq, 128);
    const int64_t nq_pad = gx * 128;
    // queries: split into hi/lo (padded to nq_pad rows of dp; padding rows are zero)
    L2Set &qs = wk.qsplit;
    qs.dp = dp;
    __bf16 *Qh = (__bf16 *)qs.hi.ensure((size_t)nq_pad * dp * 2);
    __bf16 *Ql = (__bf16 *)qs.lo.ensure((size_t)nq_pad * dp * 2);
    float *qn = (float *)qs.norm.ensure((size_t)nq_pad * 4);
    hipLaunchKernelGGL(l2_split_kernel, dim3((unsigned)ceil_div(nq_pad, 4)), dim3(256), 0, st, Q, nq, nq_pad, d,
                       qstride, dp, Qh, Ql, qn);
    // slices: whole multiples of the CU count, at least 128 rows each
    const int64_t wpc = std::max<int64_t>(1, tune_get("l2.wg_per_cu"));
    int64_t S = std::max<int64_t>(1, wpc * cus / gx);
    S = std::min<int64_t>(S, std::max<int64_t>(1, n / 128));
    const int64_t rps = ceil_div(ceil_div(n, S), 32) * 32; // slices start on 32-row boundaries
    S = ceil_div(n, rps);
    const int64_t nlists = S;
    uint64_t *part = (uint64_t *)wk.part.ensure((size_t)nlists * KL * nq_pad * 8);
    const int prec = (int)tune_get("l2.mfma_terms");
    {
        StageScope sc(st, STAGE_SCAN);
        dim3 grid((unsigned)gx, (unsigned)S);
        auto launch = [&](auto klc, auto pc) {
            constexpr int KLV = decltype(klc)::value, PV = decltype(pc)::value;
            const __bf16 *xh = xs.hi.as<__bf16>(), *xl = xs.lo.as<__bf16>();
            const float *xnr = xs.norm.as<float>();
            switch (dp) {
            case 16:
                hipLaunchKernelGGL((l2_mfma_scan_small_kernel<KLV, PV, 1>), grid, dim3(256), 0, st, xh, xl, xnr, n, Qh,
                                   Ql, qn, rps, part, nq_pad);
                break;
            case 32:
                hipLaunchKernelGGL((l2_mfma_scan_small_kernel<KLV, PV, 2>), grid, dim3(256), 0, st, xh, xl, xnr, n, Qh,
                                   Ql, qn, rps, part, nq_pad);
                break;
            case 48:
                hipLaunchKernelGGL((l2_mfma_scan_small_kernel<KLV, PV, 3>), grid, dim3(256), 0, st, xh, xl, xnr, n, Qh,
                                   Ql, qn, rps, part, nq_pad);
                break;
            case 64:
                hipLaunchKernelGGL((l2_mfma_scan_small_kernel<KLV, PV, 4>), grid, dim3(256), 0, st, xh, xl, xnr, n, Qh,
                                   Ql, qn, rps, part, nq_pad);
                break;
            default:
                hipLaunchKernelGGL((l2_mfma_scan_kernel<KLV, PV>), grid, dim3(256), 0, st, xh, xl, xnr, n, dp, Qh, Ql, qn,
                                   rps, part, nq_pad);
            }
        };
        using I8 = std::integral_constant<int, 8>;
        using I16 = std::integral_constant<int, 16>;
        using I32 = std::integral_constant<int, 32>;
        using P1 = std::integral_constant<int, 1>;
        using P3 = std::integral_constant<int, 3>;
        if (prec == 1) {
            if (KL == 8) launch(I8{}, P1{}); else if (KL == 16) launch(I16{}, P1{}); else launch(I32{}, P1{});
        } else {
            if (KL == 8) launch(I8{}, P3{}); else if (KL == 16) launch(I16{}, P3{}); else launch(I32{}, P3{});
        }
        PN_HIP(hipGetLastError());
    }
    int32_t *cidx = (int32_t *)wk.cidx.ensure((size_t)nq * K2c * 4);
    float *capx = (float *)wk.capx.ensure((size_t)nq * K2c * 4);
    int32_t *flags = (int32_t *)wk.flags.ensure((size_t)nq * 4);
    {
        StageScope sc(st, STAGE_SELECT);
        merge_lists_wave(part, nlists, KL, nq, nq_pad, K2c, OUT_F32, cidx, capx, st);
    }
    {
        StageScope sc(st, STAGE_RERANK);
        const float eps_coef = (prec == 1 ? 0.0078125f : 6.103515625e-05f) + (float)dp * 4.76837158203125e-07f;
        hipLaunchKernelGGL(l2_rerank_kernel, dim3((unsigned)ceil_div(nq, 4)), dim3(256), 0, st, X, xstride, d, Q,
                           qstride, qn, xs.norm_max.as<unsigned int>(), cidx, capx, K2c, part, nlists, KL, nq_pad, nq,
                           kk, n, eps_coef, out_idx, out_dist, flags);
        PN_HIP(hipGetLastError());
    }
    {
        StageScope sc(st, STAGE_OTHER);
#define PN_L2FB(KV)                                                                                              \
    hipLaunchKernelGGL((l2_fallback_kernel<KV>), dim3((unsigned)nq), dim3(256), 0, st, X, xstride, n, d, Q, qstride, \
                       flags, kk, out_idx, out_dist)
        if (kk <= 8) PN_L2FB(8); else if (kk <= 16) PN_L2FB(16); else PN_L2FB(32);
#undef PN_L2FB
        PN_HIP(hipGetLastError());
    }
}

} // namespace pn
