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
#include "l2_mfma_dev.hpp"
#include "ref_arith.hpp"

namespace pn {


// ---------------------------------------------------------------- single-index kernels
__global__ __launch_bounds__(256) void l2_split_kernel(const float *__restrict__ X, int64_t n, int64_t n_out,
                                                       int64_t d, int64_t stride, int dp, __bf16 *__restrict__ H,
                                                       __bf16 *__restrict__ Lo, float *__restrict__ nrm) {
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row < n_out) split_row(X, n, d, stride, dp, H, Lo, nrm, row);
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

// A workgroup owns 128 queries (block blockIdx.x) and one slice of rows (blockIdx.y).
template <int KL, int PREC>
__global__ __launch_bounds__(256) void l2_mfma_scan_kernel(const __bf16 *__restrict__ Xh, const __bf16 *__restrict__ Xl,
                                                           const float *__restrict__ xn, int64_t n, int dp,
                                                           const __bf16 *__restrict__ Qh, const __bf16 *__restrict__ Ql,
                                                           const float *__restrict__ qn, int64_t rows_per_slice,
                                                           uint64_t *__restrict__ part, int64_t nq_pad) {
    __shared__ uint64_t lds[4][32][KL];
    scan_generic<KL, PREC>(lds, Xh, Xl, xn, n, dp, Qh, Ql, qn, rows_per_slice, part, nq_pad, blockIdx.x, blockIdx.y);
}

template <int KL, int PREC, int NS>
__global__ __launch_bounds__(256) void l2_mfma_scan_small_kernel(
    const __bf16 *__restrict__ Xh, const __bf16 *__restrict__ Xl, const float *__restrict__ xn, int64_t n,
    const __bf16 *__restrict__ Qh, const __bf16 *__restrict__ Ql, const float *__restrict__ qn,
    int64_t rows_per_slice, uint64_t *__restrict__ part, int64_t nq_pad) {
    __shared__ uint64_t lds[4][32][KL];
    scan_small<KL, PREC, NS>(lds, Xh, Xl, xn, n, Qh, Ql, qn, rows_per_slice, part, nq_pad, blockIdx.x, blockIdx.y);
}

__global__ __launch_bounds__(256) void l2_rerank_kernel(const float *__restrict__ X, int64_t xstride, int64_t d,
                                                        const float *__restrict__ Q, int64_t qstride,
                                                        const float *__restrict__ qn, const unsigned int *__restrict__ xn_max_bits,
                                                        const int32_t *__restrict__ cidx, const float *__restrict__ capx,
                                                        int K2c, const uint64_t *__restrict__ part, int64_t nlists,
                                                        int KL, int64_t nq_pad, int64_t nq, int kk, int64_t n,
                                                        float eps_coef, int32_t *__restrict__ out_idx,
                                                        float *__restrict__ out_dist, int32_t *__restrict__ flags) {
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q < nq)
        rerank_query(X, xstride, d, Q, qstride, qn, xn_max_bits, cidx, capx, K2c, part, nlists, KL, nq_pad, kk, n,
                     eps_coef, out_idx, out_dist, flags, q);
}

// One workgroup per query; unflagged queries exit at once.
template <int K>
__global__ __launch_bounds__(256) void l2_fallback_kernel(const float *__restrict__ X, int64_t xstride, int64_t n,
                                                          int64_t d, const float *__restrict__ Q, int64_t qstride,
                                                          const int32_t *__restrict__ flags, int kk,
                                                          int32_t *__restrict__ out_idx, float *__restrict__ out_dist) {
    __shared__ uint64_t wres[4][K];
    if (flags[blockIdx.x]) fallback_query<K>(wres, X, xstride, n, d, Q, qstride, kk, out_idx, out_dist, blockIdx.x);
}

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
        merge_lists_wave(part, nlists, KL, K2, nq, nq_pad, K2c, OUT_F32, cidx, capx, st); // capacity >= K2c
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
