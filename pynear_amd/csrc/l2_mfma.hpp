// Host interface of the MFMA L2 search (l2_mfma.hip).
#pragma once
#include "common.hpp"

namespace pn {

// bf16 hi/lo split of a row set with fp32 squared norms and the largest norm.
struct L2Set {
    DevBuf hi, lo, norm, norm_max;
    int64_t n = 0, d = 0;
    int dp = 0; // d rounded up to 16
};

// Per-index scratch of a search.
struct L2Work {
    L2Set qsplit;
    DevBuf part, cidx, capx, flags;
};

// X: device rows (fp32, row stride `stride` floats).
void l2_prepare(L2Set &s, const float *X, int64_t n, int64_t d, int64_t stride, hipStream_t st);

// bf16 hi/lo split + squared norms of n rows into caller buffers (rows of dp elements).
void l2_split_rows(const float *X, int64_t n, int64_t d, int64_t stride, int dp, __bf16 *H, __bf16 *L, float *nrm,
                   hipStream_t st);

// Exact k-NN (kk <= 32) of nq device queries Q (row stride qstride) against the rows X the split
// `xs` was prepared from; out_idx int32[nq, kk], out_dist float32[nq, kk] (sqrt distances),
// ascending (distance, index), bit-identical to the reference's dist_l2_f_avx2.
void l2_mfma_search(const L2Set &xs, const float *X, int64_t xstride, const float *Q, int64_t qstride, int64_t nq,
                    int kk, int32_t *out_idx, float *out_dist, L2Work &wk, int cus, hipStream_t st);

} // namespace pn
