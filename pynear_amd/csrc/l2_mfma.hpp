// Host interface of the MFMA L2 search (l2_mfma.hip).
#pragma once
#include "common.hpp"

namespace pn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// acc += X-tile . Q-tile over dp dims (multiple of 16) with v_mfma_f32_32x32x16_bf16: PREC 1 uses
// hi.hi, PREC 3 adds hi.lo + lo.hi. The loads of CH consecutive K steps are issued together before
// their products, so a tile pays one memory round trip per CH steps instead of one per step.
template <int PREC, int CH>
__device__ __forceinline__ void mfma_tile_bf16(f32x16 &acc, const __bf16 *xh, const __bf16 *xl, const __bf16 *qh,
                                               const __bf16 *ql, int dp) {
    int kk = 0;
    for (; kk + 16 * CH <= dp; kk += 16 * CH) {
        bf16x8 a[CH], b[CH], al[CH], bl[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            a[u] = *(const bf16x8 *)(xh + kk + 16 * u);
            b[u] = *(const bf16x8 *)(qh + kk + 16 * u);
            if (PREC == 3) {
                al[u] = *(const bf16x8 *)(xl + kk + 16 * u);
                bl[u] = *(const bf16x8 *)(ql + kk + 16 * u);
            }
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[u], b[u], acc, 0, 0, 0);
            if (PREC == 3) {
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[u], bl[u], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[u], b[u], acc, 0, 0, 0);
            }
        }
    }
    for (; kk < dp; kk += 16) {
        const bf16x8 a = *(const bf16x8 *)(xh + kk), b = *(const bf16x8 *)(qh + kk);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        if (PREC == 3) {
            const bf16x8 al = *(const bf16x8 *)(xl + kk), bl = *(const bf16x8 *)(ql + kk);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bl, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, b, acc, 0, 0, 0);
        }
    }
}

// bf16 hi/lo split of a row set with fp32 squared norms and the largest norm.
struct L2Set {
    DevBuf hi, lo, norm, norm_max;
    int64_t n = 0, d = 0;
    int dp = 0; // d rounded up to the alignment l2_prepare was given (16 by default)
};

// Per-index scratch of a search.
struct L2Work {
    L2Set qsplit;
    DevBuf part, cidx, capx, flags;
};

// X: device rows (fp32, row stride `stride` floats).
void l2_prepare(L2Set &s, const float *X, int64_t n, int64_t d, int64_t stride, hipStream_t st, int align = 16);

// bf16 hi/lo split + squared norms of n rows into caller buffers (rows of dp elements).
void l2_split_rows(const float *X, int64_t n, int64_t d, int64_t stride, int dp, __bf16 *H, __bf16 *L, float *nrm,
                   hipStream_t st);

// Exact k-NN (kk <= 32) of nq device queries Q (row stride qstride) against the rows X the split
// `xs` was prepared from; out_idx int32[nq, kk], out_dist float32[nq, kk] (sqrt distances),
// ascending (distance, index), bit-identical to the reference's dist_l2_f_avx2.
void l2_mfma_search(const L2Set &xs, const float *X, int64_t xstride, const float *Q, int64_t qstride, int64_t nq,
                    int kk, int32_t *out_idx, float *out_dist, L2Work &wk, int cus, hipStream_t st);

} // namespace pn
