// Device restatement of the reference's float distance arithmetic, operation for operation, so
// that distances are bit-identical to DistanceFunctions.hpp:
//   L2  (dist_l2_f_avx2, :304-345): 8 FMA lanes over full blocks of 8, high half folded onto the
//       low half, a fused 4-lane block and a fused masked 1..3 tail (GCC contracts the tail's
//       mul+add with -mfma), two horizontal adds, sqrt. l2_core() stops before the sqrt.
//   L1  (dist_l1_f_avx2, :347-382): 8 lane sums of |a-b|, summed lane 0..7, then a scalar tail.
//   Chebyshev (:384-420): max |a-b| (exact, order-free).
// sqrtf is correctly rounded (HIP's default fp32 sqrt), like std::sqrt.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/pynear_amd.h"

namespace pn {

__device__ inline float l2_core(const float *a, const float *b, int64_t d) {
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int64_t i = 0;
    for (; d - i >= 8; i += 8) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            float t = a[i + j] - b[i + j];
            acc[j] = fmaf(t, t, acc[j]);
        }
    }
    float m[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) m[j] = acc[4 + j] + acc[j];
    for (int pass = 0; pass < 2; ++pass) {
        int64_t rem = d - i;
        int w = pass == 0 ? (rem >= 4 ? 4 : 0) : (int)rem;
        if (w == 0) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float t = j < w ? a[i + j] - b[i + j] : 0.0f;
            m[j] = fmaf(t, t, m[j]);
        }
        i += w;
    }
    return (m[0] + m[1]) + (m[2] + m[3]);
}

// End of the 8-lane squared sum (below): lane c holds FMA chain c over the full blocks of 8; the
// fold, the fused tails and the horizontal adds follow. The result is valid in lane 0 of the group.
__device__ __forceinline__ float l2_group8_finish(float acc, const float *a, const float *b, int64_t d, int c) {
    const int64_t nb = d / 8 * 8;
    const float hi = __shfl_down(acc, 4, 8);
    float m = hi + acc; // lanes 0..3: m[c] = acc[c + 4] + acc[c]
    if (c < 4) {
        int64_t i = nb, rem = d - nb;
        if (rem >= 4) {
            const float t = a[i + c] - b[i + c];
            m = fmaf(t, t, m);
            i += 4;
            rem -= 4;
        }
        if (rem > 0) {
            const float t = c < rem ? a[i + c] - b[i + c] : 0.f;
            m = fmaf(t, t, m);
        }
    }
    const float p = m + __shfl_down(m, 1, 8); // lane 0: m0 + m1; lane 2: m2 + m3
    return p + __shfl_down(p, 2, 8);
}

// Squared-sum core of dist_l2_f_avx2 with 8 lanes per row: lane c runs FMA chain c (elements
// c, c+8, ...), then the fold, the fused tails and the horizontal adds; the result is valid in
// lane 0 of the group.
__device__ __forceinline__ float l2_core_group8(const float *a, const float *b, int64_t d, int c) {
    const int64_t nb = d / 8 * 8;
    float acc = 0.f;
    int64_t i = c;
    for (; i + 56 < nb; i += 64) { // the loads of 8 steps first, then the chain in order
        float ta[8], tb[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            ta[u] = a[i + 8 * u];
            tb[u] = b[i + 8 * u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const float t = ta[u] - tb[u];
            acc = fmaf(t, t, acc);
        }
    }
    for (; i < nb; i += 8) {
        const float t = a[i] - b[i];
        acc = fmaf(t, t, acc);
    }
    return l2_group8_finish(acc, a, b, d, c);
}

__device__ inline float l1_dist(const float *a, const float *b, int64_t d) {
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int64_t i = 0;
    for (; i + 8 <= d; i += 8) {
#pragma unroll
        for (int j = 0; j < 8; ++j) s[j] = s[j] + fabsf(a[i + j] - b[i + j]);
    }
    float total = 0.f;
#pragma unroll
    for (int j = 0; j < 8; ++j) total += s[j];
    for (; i < d; ++i) total += fabsf(a[i] - b[i]);
    return total;
}

__device__ inline float linf_dist(const float *a, const float *b, int64_t d) {
    float m = 0.f;
    for (int64_t i = 0; i < d; ++i) m = fmaxf(m, fabsf(a[i] - b[i]));
    return m;
}

// Final reference distance for a float metric.
__device__ inline float ref_distance(int metric, const float *a, const float *b, int64_t d) {
    if (metric == PN_METRIC_L1) return l1_dist(a, b, d);
    if (metric == PN_METRIC_CHEBYSHEV) return linf_dist(a, b, d);
    return sqrtf(l2_core(a, b, d));
}

} // namespace pn
