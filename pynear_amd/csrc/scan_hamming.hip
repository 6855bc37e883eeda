// Exact Hamming k-NN scan: popcount(q xor x) over packed rows, per-lane top-K lists.
//
// Replaces the per-node dist_hamming_* calls of the VP-tree walk (VPTree.hpp:437-508 with
// DistanceFunctions.hpp:166-169, :431-462) by a brute-force scan of every row.
//
// Mapping (one wave = 64 queries, one per lane; the database is streamed through SGPRs):
//   * a lane keeps its query's W 32-bit words in VGPRs;
//   * every row of the slice is wave-uniform, so the compiler fetches it with s_load_dwordx16
//     and each word costs one v_xor_b32 (SGPR operand) + one v_bcnt_u32_b32 (popcount-accumulate);
//   * the lane's top-K list lives in VGPRs as packed (d << 21 | row-in-slice) keys. Rows arrive in
//     increasing index order, so a new row only enters when d < d(L[K-1]) (an equal distance
//     loses the index tie); the insert is K v_med3_u32 and is skipped by the whole wave when no
//     lane qualifies, which is the common case once the lists fill.
//   * slices of the database run in separate workgroups (gridDim.y); their lists are merged by
//     the select kernel.
#include "common.hpp"
#include "device_util.hpp"
#include "kernels.hpp"

namespace pn {

template <int W, int K, int RU>
__global__ __launch_bounds__(256) void hamming_scan_kernel(const uint32_t *__restrict__ db, int64_t n,
                                                           int64_t rows_per_slice, const uint32_t *__restrict__ qs,
                                                           int64_t nq, uint64_t *__restrict__ part, int64_t nq_pad) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t s = blockIdx.y;
    const int64_t r0 = s * rows_per_slice;
    const int64_t r1 = min(n, r0 + rows_per_slice);

    uint32_t qv[W];
    if (q < nq) {
#pragma unroll
        for (int w = 0; w < W; ++w) qv[w] = qs[q * W + w];
    } else {
#pragma unroll
        for (int w = 0; w < W; ++w) qv[w] = 0u;
    }
    uint32_t L[K];
#pragma unroll
    for (int j = 0; j < K; ++j) L[j] = kKey32None;
    uint32_t thr = kKey32None >> kHamIdxBits;

    int64_t r = r0;
    for (; r + RU <= r1; r += RU) {
        uint32_t d[RU];
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const uint32_t *row = db + (r + u) * W;
            uint32_t acc = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) acc = __popc(qv[w] ^ row[w]) + acc;
            d[u] = acc;
        }
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            if (d[u] < thr) {
                list_insert_u32<K>(L, (d[u] << kHamIdxBits) | (uint32_t)(r + u - r0));
                thr = L[K - 1] >> kHamIdxBits;
            }
        }
    }
    for (; r < r1; ++r) {
        const uint32_t *row = db + r * W;
        uint32_t acc = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) acc = __popc(qv[w] ^ row[w]) + acc;
        if (acc < thr) {
            list_insert_u32<K>(L, (acc << kHamIdxBits) | (uint32_t)(r - r0));
            thr = L[K - 1] >> kHamIdxBits;
        }
    }

    if (q < nq) {
        uint64_t *out = part + (s * K) * nq_pad + q;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            uint32_t key = L[j];
            uint64_t g = kKey64None;
            if (key != kKey32None)
                g = ((uint64_t)(key >> kHamIdxBits) << 32) | (uint64_t)(r0 + (key & kHamIdxMask));
            out[(int64_t)j * nq_pad] = g;
        }
    }
}

template <int W, int K>
static void launch_w_k(const ScanArgs &a, hipStream_t st) {
    dim3 grid((unsigned)ceil_div(a.nq, 256), (unsigned)a.slices);
    hipLaunchKernelGGL((hamming_scan_kernel<W, K, 4>), grid, dim3(256), 0, st, (const uint32_t *)a.db, a.n,
                       a.rows_per_slice, (const uint32_t *)a.queries, a.nq, a.part, a.nq_pad);
}

template <int W> static void launch_w(const ScanArgs &a, int K, hipStream_t st) {
    switch (K) {
    case 1: launch_w_k<W, 1>(a, st); break;
    case 2: launch_w_k<W, 2>(a, st); break;
    case 4: launch_w_k<W, 4>(a, st); break;
    case 8: launch_w_k<W, 8>(a, st); break;
    case 16: launch_w_k<W, 16>(a, st); break;
    case 32: launch_w_k<W, 32>(a, st); break;
    default: fail(PN_EINVAL, "hamming scan: unsupported list size");
    }
}

void hamming_scan(const ScanArgs &a, int words, int K, hipStream_t st) {
    PN_REQUIRE(a.rows_per_slice <= kMaxRowsPerSlice, PN_EINVAL, "hamming scan: slice too long");
    switch (words) {
    case 1: launch_w<1>(a, K, st); break;
    case 2: launch_w<2>(a, K, st); break;
    case 4: launch_w<4>(a, K, st); break;
    case 8: launch_w<8>(a, K, st); break;
    case 16: launch_w<16>(a, K, st); break;
    case 32: launch_w<32>(a, K, st); break;
    default: fail(PN_EINVAL, "hamming scan: unsupported row width");
    }
    PN_HIP(hipGetLastError());
}

int hamming_padded_words(int64_t nbytes) {
    int64_t w = ceil_div(nbytes, 4);
    for (int c : {1, 2, 4, 8, 16, 32})
        if (w <= c) return c;
    return 0; // wider rows take the distance-matrix path
}

} // namespace pn
