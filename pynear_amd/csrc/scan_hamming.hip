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
//   * G > 0 (multi-index hashing, MIH.hpp:98-150): the row's sub-strings sit in groups of G words;
//     a row enters only when one of its first `subs` groups is within `rsub` bits of the query's
//     (the candidate set of the sub-string tables), with the full distance as its key.
#include <algorithm>

#include "common.hpp"
#include "device_util.hpp"
#include "kernels.hpp"

namespace pn {

template <int W, int K, int QPL, int RU, int G = 0>
__global__ __launch_bounds__(256) void hamming_scan_kernel(const uint32_t *__restrict__ db, int64_t n,
                                                           int64_t rows_per_slice, const uint32_t *__restrict__ qs,
                                                           int64_t nq, uint64_t *__restrict__ part, int64_t nq_pad,
                                                           int subs = 0, uint32_t rsub = 0) {
    // lane t of the workgroup serves queries base + t + 256 p, p < QPL (each wave stays dense)
    const int64_t qbase = (int64_t)blockIdx.x * (256 * QPL) + threadIdx.x;
    const int64_t s = blockIdx.y;
    const int64_t r0 = s * rows_per_slice;
    const int64_t r1 = min(n, r0 + rows_per_slice);

    uint32_t qv[QPL][W];
#pragma unroll
    for (int p = 0; p < QPL; ++p) {
        const int64_t q = qbase + 256 * p;
#pragma unroll
        for (int w = 0; w < W; ++w) qv[p][w] = q < nq ? qs[q * W + w] : 0u;
    }
    uint32_t L[QPL][K], thr[QPL];
#pragma unroll
    for (int p = 0; p < QPL; ++p) {
#pragma unroll
        for (int j = 0; j < K; ++j) L[p][j] = kKey32None;
        thr[p] = kKey32None >> kHamIdxBits;
    }

    // Rows are wave-uniform: the loads below become s_load_dwordx16 into SGPRs (scalar loads
    // complete out of order, so each group of RU rows is waited for as a whole; QPL queries per
    // lane give every loaded row QPL x 2W VALU operations). Rows past the slice end are clamped to
    // its last row (valid memory) and never inserted.
    // 32-bit row arithmetic inside the slice keeps the clamp on the scalar unit (s_min_i32).
    const uint32_t *sdb = db + r0 * W;
    const int nrows = (int)(r1 - r0);
    auto load_rows = [&](uint32_t (&R)[RU][W], int lr) {
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const int rr = min(lr + u, nrows - 1);
#pragma unroll
            for (int w = 0; w < W; ++w) R[u][w] = sdb[rr * W + w];
        }
    };
    auto compute = [&](const uint32_t (&R)[RU][W], int lr) {
        const int cnt = min(RU, nrows - lr);
        uint32_t d[QPL][RU];
#pragma unroll
        for (int u = 0; u < RU; ++u)
#pragma unroll
            for (int p = 0; p < QPL; ++p) {
                if constexpr (G == 0) {
                    uint32_t acc = 0;
#pragma unroll
                    for (int w = 0; w < W; ++w) acc = bcnt_acc(qv[p][w] ^ R[u][w], acc);
                    d[p][u] = acc;
                } else {
                    uint32_t acc = 0, mn = kKey32None;
#pragma unroll
                    for (int g = 0; g < W / G; ++g) {
                        uint32_t s = 0;
#pragma unroll
                        for (int j = 0; j < G; ++j) s = bcnt_acc(qv[p][g * G + j] ^ R[u][g * G + j], s);
                        acc += s;
                        mn = g < subs ? min(mn, s) : mn;
                    }
                    d[p][u] = mn <= rsub ? acc : kKey32None; // a non-candidate never passes thr
                }
            }
#pragma unroll
        for (int p = 0; p < QPL; ++p) {
            uint32_t mn = d[p][0];
#pragma unroll
            for (int u = 1; u < RU; ++u) mn = min(mn, d[p][u]);
            if (mn < thr[p]) {
#pragma unroll
                for (int u = 0; u < RU; ++u) {
                    if (u < cnt && d[p][u] < thr[p]) {
                        list_insert_u32<K>(L[p], (d[p][u] << kHamIdxBits) | (uint32_t)(lr + u));
                        thr[p] = L[p][K - 1] >> kHamIdxBits;
                    }
                }
            }
        }
    };

    uint32_t R[RU][W];
    for (int lr = 0; lr < nrows; lr += RU) {
        load_rows(R, lr);
        __builtin_amdgcn_sched_barrier(0); // all RU row loads in flight before the first use
        compute(R, lr);
    }

#pragma unroll
    for (int p = 0; p < QPL; ++p) {
        const int64_t q = qbase + 256 * p;
        if (q < nq) {
            uint64_t *out = part + (s * K) * nq_pad + q;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                uint32_t key = L[p][j];
                uint64_t g = kKey64None;
                if (key != kKey32None)
                    g = ((uint64_t)(key >> kHamIdxBits) << 32) | (uint64_t)(r0 + (key & kHamIdxMask));
                out[(int64_t)j * nq_pad] = g;
            }
        }
    }
}

template <int W, int K>
static void launch_w_k(const ScanArgs &a, hipStream_t st) {
    dim3 grid((unsigned)ceil_div(a.nq, 256 * a.qpl), (unsigned)a.slices);
    // SGPR budget: RU x W row words must fit next to the addressing registers.
    constexpr int RU = W >= 32 ? 2 : (W >= 16 ? 4 : 8);
    if (a.filt.group_words > 0) {
        PN_REQUIRE(a.qpl == 1, PN_EINVAL, "hamming scan: the candidate filter takes one query per lane");
        const uint32_t rs = (uint32_t)std::max(0, a.filt.rsub);
        if constexpr (K != 4 && K != 16 && K != 32)
            fail(PN_EINVAL, "hamming scan: the candidate filter takes lists of 4, 16 or 32");
        else if (a.filt.group_words == 1)
            hipLaunchKernelGGL((hamming_scan_kernel<W, K, 1, RU, 1>), grid, dim3(256), 0, st, (const uint32_t *)a.db,
                               a.n, a.rows_per_slice, (const uint32_t *)a.queries, a.nq, a.part, a.nq_pad, a.filt.subs,
                               rs);
        else if constexpr (W >= 2)
            hipLaunchKernelGGL((hamming_scan_kernel<W, K, 1, RU, 2>), grid, dim3(256), 0, st, (const uint32_t *)a.db,
                               a.n, a.rows_per_slice, (const uint32_t *)a.queries, a.nq, a.part, a.nq_pad, a.filt.subs,
                               rs);
        else
            fail(PN_EINVAL, "hamming scan: groups of two words need rows of two words or more");
        return;
    }
    if (a.qpl == 2)
        hipLaunchKernelGGL((hamming_scan_kernel<W, K, 2, RU>), grid, dim3(256), 0, st, (const uint32_t *)a.db, a.n,
                           a.rows_per_slice, (const uint32_t *)a.queries, a.nq, a.part, a.nq_pad);
    else
        hipLaunchKernelGGL((hamming_scan_kernel<W, K, 1, RU>), grid, dim3(256), 0, st, (const uint32_t *)a.db, a.n,
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
