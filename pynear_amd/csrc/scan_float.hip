// Exact L2 / L1 / Chebyshev k-NN scan for rows of up to 64 floats, in the reference's arithmetic.
//
// Replaces the per-node dist_*_f_avx2 calls of the VP-tree walk (VPTree.hpp:437-508 with
// DistanceFunctions.hpp:304-420) by a brute-force scan of every row. Distances are computed with
// the reference's exact operation order (see ref_arith.hpp), so every distance is bit-identical to
// the reference's.
//
// Mapping (one wave = 64 queries, one per lane; rows streamed through SGPRs): a lane holds its
// query padded to P floats in VGPRs; each row of the slice is wave-uniform (s_load), so each
// element costs v_sub (SGPR operand) + v_fma (or v_add/v_max with |x|). Rows are padded with
// zeros to P floats on upload; zero lanes are exact no-ops in every metric, which lets the
// reference's masked tail run unconditionally. The lane keeps a top-K list of (value, row)
// where value is the pre-sqrt sum for L2 (sqrt is monotone; it is applied in the select stage).
#include "common.hpp"
#include "device_util.hpp"
#include "kernels.hpp"

namespace pn {

template <int P, int K, int METRIC, int RU>
__global__ __launch_bounds__(256) void float_scan_kernel(const float *__restrict__ db, int64_t n, int64_t dim,
                                                         int64_t rows_per_slice, const float *__restrict__ qs,
                                                         int64_t nq, uint64_t *__restrict__ part, int64_t nq_pad) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t s = blockIdx.y;
    const int64_t r0 = s * rows_per_slice;
    const int64_t r1 = min(n, r0 + rows_per_slice);
    const int nb = (int)(dim / 8);          // full 8-blocks
    const bool tail = (dim % 8) != 0;

    float qv[P];
    if (q < nq) {
#pragma unroll
        for (int e = 0; e < P; ++e) qv[e] = qs[q * P + e];
    } else {
#pragma unroll
        for (int e = 0; e < P; ++e) qv[e] = 0.f;
    }
    uint32_t Lk[K], Li[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        Lk[j] = kKey32None;
        Li[j] = 0u;
    }
    uint32_t thr = kKey32None;

    auto row_value = [&](const float *row) -> float {
        if (METRIC == PN_METRIC_L2) {
            float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int blk = 0; blk < P / 8; ++blk) {
                if (blk < nb) {
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        float t = qv[blk * 8 + j] - row[blk * 8 + j];
                        acc[j] = fmaf(t, t, acc[j]);
                    }
                }
            }
            float m[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) m[j] = acc[4 + j] + acc[j];
            if (tail) {
#pragma unroll
                for (int blk = 0; blk < P / 8; ++blk) {
                    if (blk == nb) {
#pragma unroll
                        for (int h = 0; h < 2; ++h)
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                float t = qv[blk * 8 + 4 * h + j] - row[blk * 8 + 4 * h + j];
                                m[j] = fmaf(t, t, m[j]);
                            }
                    }
                }
            }
            return (m[0] + m[1]) + (m[2] + m[3]);
        } else if (METRIC == PN_METRIC_L1) {
            float sl[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int blk = 0; blk < P / 8; ++blk) {
                if (blk < nb) {
#pragma unroll
                    for (int j = 0; j < 8; ++j) sl[j] = sl[j] + fabsf(qv[blk * 8 + j] - row[blk * 8 + j]);
                }
            }
            float total = 0.f;
#pragma unroll
            for (int j = 0; j < 8; ++j) total += sl[j];
            if (tail) {
#pragma unroll
                for (int blk = 0; blk < P / 8; ++blk) {
                    if (blk == nb) {
#pragma unroll
                        for (int j = 0; j < 8; ++j) total += fabsf(qv[blk * 8 + j] - row[blk * 8 + j]);
                    }
                }
            }
            return total;
        } else {
            float mx = 0.f;
#pragma unroll
            for (int e = 0; e < P; ++e) mx = fmaxf(mx, fabsf(qv[e] - row[e]));
            return mx;
        }
    };

    auto consider = [&](uint32_t key, uint32_t rl) {
        if (key < thr) {
#pragma unroll
            for (int j = K - 1; j > 0; --j) {
                bool up = key < Lk[j - 1];
                bool here = key < Lk[j];
                Li[j] = up ? Li[j - 1] : (here ? rl : Li[j]);
                Lk[j] = max(Lk[j - 1], min(key, Lk[j]));
            }
            if (key < Lk[0]) {
                Li[0] = rl;
                Lk[0] = key;
            }
            thr = Lk[K - 1];
        }
    };

    int64_t r = r0;
    for (; r + RU <= r1; r += RU) {
        uint32_t v[RU];
#pragma unroll
        for (int u = 0; u < RU; ++u) v[u] = f32_key(row_value(db + (r + u) * P));
#pragma unroll
        for (int u = 0; u < RU; ++u) consider(v[u], (uint32_t)(r + u - r0));
    }
    for (; r < r1; ++r) consider(f32_key(row_value(db + r * P)), (uint32_t)(r - r0));

    if (q < nq) {
        uint64_t *out = part + (s * K) * nq_pad + q;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            uint64_t g = kKey64None;
            if (Lk[j] != kKey32None) g = ((uint64_t)Lk[j] << 32) | (uint64_t)(r0 + Li[j]);
            out[(int64_t)j * nq_pad] = g;
        }
    }
}

template <int P, int K, int M> static void launch_pkm(const ScanArgs &a, hipStream_t st) {
    dim3 grid((unsigned)ceil_div(a.nq, 256), (unsigned)a.slices);
    hipLaunchKernelGGL((float_scan_kernel<P, K, M, 2>), grid, dim3(256), 0, st, (const float *)a.db, a.n, a.dim,
                       a.rows_per_slice, (const float *)a.queries, a.nq, a.part, a.nq_pad);
}

template <int P, int M> static void launch_pm(const ScanArgs &a, int K, hipStream_t st) {
    switch (K) {
    case 1: launch_pkm<P, 1, M>(a, st); break;
    case 2: launch_pkm<P, 2, M>(a, st); break;
    case 4: launch_pkm<P, 4, M>(a, st); break;
    case 8: launch_pkm<P, 8, M>(a, st); break;
    case 16: launch_pkm<P, 16, M>(a, st); break;
    case 32: launch_pkm<P, 32, M>(a, st); break;
    default: fail(PN_EINVAL, "float scan: unsupported list size");
    }
}

template <int P> static void launch_p(const ScanArgs &a, int metric, int K, hipStream_t st) {
    switch (metric) {
    case PN_METRIC_L2: launch_pm<P, PN_METRIC_L2>(a, K, st); break;
    case PN_METRIC_L1: launch_pm<P, PN_METRIC_L1>(a, K, st); break;
    case PN_METRIC_CHEBYSHEV: launch_pm<P, PN_METRIC_CHEBYSHEV>(a, K, st); break;
    default: fail(PN_EINVAL, "float scan: bad metric");
    }
}

void float_scan(const ScanArgs &a, int metric, int K, hipStream_t st) {
    PN_REQUIRE(a.rows_per_slice <= (int64_t)0xFFFFFFFF, PN_EINVAL, "float scan: slice too long");
    switch (a.row_elems) {
    case 8: launch_p<8>(a, metric, K, st); break;
    case 16: launch_p<16>(a, metric, K, st); break;
    case 32: launch_p<32>(a, metric, K, st); break;
    case 64: launch_p<64>(a, metric, K, st); break;
    default: fail(PN_EINVAL, "float scan: unsupported padded dimension");
    }
    PN_HIP(hipGetLastError());
}

} // namespace pn
