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
#include "merge_dev.hpp"
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

// A workgroup owns 128 queries (one query block) and one slice of rows: with slice_major the
// slice is blockIdx.x and the block blockIdx.y, else the other way round.
template <int KL, int PREC>
__global__ __launch_bounds__(256) void l2_mfma_scan_kernel(const __bf16 *__restrict__ Xh, const __bf16 *__restrict__ Xl,
                                                           const float *__restrict__ xn, int64_t n, int dp,
                                                           const __bf16 *__restrict__ Qh, const __bf16 *__restrict__ Ql,
                                                           const float *__restrict__ qn, int64_t rows_per_slice,
                                                           uint64_t *__restrict__ part, int64_t qsl, bool slice_major) {
    __shared__ uint64_t lds[4][32][KL];
    scan_generic<KL, PREC>(lds, Xh, Xl, xn, n, dp, Qh, Ql, qn, rows_per_slice, part, qsl,
                           slice_major ? blockIdx.y : blockIdx.x, slice_major ? blockIdx.x : blockIdx.y);
}

template <int KL, int PREC, int NS>
__global__ __launch_bounds__(256) void l2_mfma_scan_small_kernel(
    const __bf16 *__restrict__ Xh, const __bf16 *__restrict__ Xl, const float *__restrict__ xn, int64_t n,
    const __bf16 *__restrict__ Qh, const __bf16 *__restrict__ Ql, const float *__restrict__ qn,
    int64_t rows_per_slice, uint64_t *__restrict__ part, int64_t qsl, bool slice_major) {
    __shared__ uint64_t lds[4][32][KL];
    // slice_major: blockIdx.x is the slice, so workgroup b (on XCD b % 8) of every query block
    // scans the same eighth of the slices, and those rows stay in that XCD's L2
    scan_small<KL, PREC, NS>(lds, Xh, Xl, xn, n, Qh, Ql, qn, rows_per_slice, part, qsl,
                             slice_major ? blockIdx.y : blockIdx.x, slice_major ? blockIdx.x : blockIdx.y);
}

template <int KL, int PREC, int NC>
__global__ __launch_bounds__(256) void l2_mfma_scan_wide_kernel(
    const __bf16 *__restrict__ Xh, const __bf16 *__restrict__ Xl, const float *__restrict__ xn, int64_t n,
    const __bf16 *__restrict__ Qh, const __bf16 *__restrict__ Ql, const float *__restrict__ qn,
    int64_t rows_per_slice, uint64_t *__restrict__ part, int64_t qsl, bool slice_major) {
    __shared__ uint64_t lds[4][32][KL];
    __shared__ __align__(16) __bf16 stage[wide_stage_elems<NC, PREC == 3 ? 2 : 1>()];
    scan_wide<KL, PREC, NC>(lds, stage, Xh, Xl, xn, n, Qh, Ql, qn, rows_per_slice, part, qsl,
                            slice_major ? blockIdx.y : blockIdx.x, slice_major ? blockIdx.x : blockIdx.y);
}

template <bool REG>
__global__ __launch_bounds__(256) void l2_rerank_kernel(const float *__restrict__ X, int64_t xstride, int64_t d,
                                                        const float *__restrict__ Q, int64_t qstride,
                                                        const float *__restrict__ qn, const unsigned int *__restrict__ xn_max_bits,
                                                        const int32_t *__restrict__ cidx, const float *__restrict__ capx,
                                                        int K2c, const uint64_t *__restrict__ part, int64_t nlists,
                                                        int KL, int64_t qsl, int64_t nq, int kk, int64_t n,
                                                        float eps_coef, int32_t *__restrict__ out_idx,
                                                        float *__restrict__ out_dist, int32_t *__restrict__ flags) {
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q < nq)
        rerank_query<REG>(X, xstride, d, Q, qstride, qn, xn_max_bits, cidx, capx, K2c, part, nlists, KL, qsl, kk, n,
                     eps_coef, out_idx, out_dist, flags, q);
}

// The wave merge of a query's slice lists (query-major, KL keys each) and the exact rerank of its
// K2c candidates in one launch: the candidates pass from the merge to the rerank in registers.
// K: the merge's private-list capacity; 0: the merge without private lists.
template <int K, bool REG>
__global__ __launch_bounds__(256) void l2_merge_rerank_kernel(
    const uint64_t *__restrict__ part, int64_t nlists, int KL, int64_t qsl, int K2c, bool rounds_only,
    const float *__restrict__ X, int64_t xstride, int64_t d, const float *__restrict__ Q, int64_t qstride,
    const float *__restrict__ qn, const unsigned int *__restrict__ xn_max_bits, int64_t nq, int kk, int64_t n,
    float eps_coef, int32_t *__restrict__ out_idx, float *__restrict__ out_dist, int32_t *__restrict__ flags) {
    __shared__ uint64_t wbuf[4][256];
    const int64_t q = xcd_contiguous_block() * 4 + (threadIdx.x >> 6);
    if (q >= nq) return; // whole wave exits together
    uint64_t mine;
    if constexpr (K == 0)
        merge_wave_lean<true>(part, nlists, KL, qsl, 1, q, K2c, OUT_F32, nullptr, nullptr, wbuf[threadIdx.x >> 6],
                              rounds_only, mine);
    else
        merge_wave_query<K, true>(part, nlists, KL, qsl, 1, q, K2c, OUT_F32, nullptr, nullptr, wbuf[threadIdx.x >> 6],
                                  mine);
    rerank_regs<REG>(X, xstride, d, Q, qstride, qn, xn_max_bits, mine, K2c, part, nlists, KL, qsl, kk, n, eps_coef,
                     out_idx, out_dist, flags, q);
}

// One workgroup per chunk of kFallbackChunk queries: their flags are read at once, and the
// flagged queries run the exact scan one after another (chunks in parallel).
template <int K>
__global__ __launch_bounds__(256) void l2_fallback_kernel(const float *__restrict__ X, int64_t xstride, int64_t n,
                                                          int64_t d, const float *__restrict__ Q, int64_t qstride,
                                                          const int32_t *__restrict__ flags, int kk,
                                                          int32_t *__restrict__ out_idx, float *__restrict__ out_dist,
                                                          int64_t nq) {
    __shared__ uint64_t wres[4][K];
    __shared__ uint32_t s_flagged;
    for (int64_t q0 = (int64_t)blockIdx.x * kFallbackChunk; q0 < nq; q0 += (int64_t)gridDim.x * kFallbackChunk) {
        if (threadIdx.x < 64) { // (wave 0: bit t for query q0 + t)
            const bool f = threadIdx.x < kFallbackChunk && q0 + threadIdx.x < nq && flags[q0 + threadIdx.x];
            const uint64_t bal = __ballot(f);
            if (threadIdx.x == 0) s_flagged = (uint32_t)bal;
        }
        __syncthreads();
        const uint32_t fl = s_flagged;
        __syncthreads(); // (s_flagged is rewritten for the next chunk)
        for (uint32_t m = fl; m; m &= m - 1)
            fallback_query<K>(wres, X, xstride, n, d, Q, qstride, kk, out_idx, out_dist, q0 + __builtin_ctz(m));
    }
}

// ---------------------------------------------------------------- small-database path
// For a few thousand rows of large d (IVF coarse centroids, k-means centroids) the approximate
// distances of a query chunk to all rows are one product (bf16, three terms per K step) written
// as a key matrix; a wave per query then takes the K2c smallest (key, row) pairs, and the exact
// rerank, proof and fallback are the scan path's. A workgroup computes a 128 x 128 block (wave:
// 64 x 64 as 2 x 2 MFMA tiles) from operands staged in LDS; rows are clamped on load and masked
// on store.
template <int PREC>
__global__ __launch_bounds__(256, 3) void l2_matrix_keys_kernel(const __bf16 *__restrict__ Qh, const __bf16 *__restrict__ Ql,
                                                             const float *__restrict__ qn, int M,
                                                             const __bf16 *__restrict__ Xh, const __bf16 *__restrict__ Xl,
                                                             const float *__restrict__ xn, int N, int dp,
                                                             uint32_t *__restrict__ keys) {
    constexpr int NPL = PREC == 3 ? 2 : 1; // planes: hi (and lo)
    constexpr int KC = 32;                  // columns per staged chunk
    constexpr int RS = KC + 8;              // LDS row stride (bf16): 80 bytes, conflict-free b128 reads
    constexpr int NB = 1;                   // chunk buffers (1: 40 KB of LDS, three workgroups per CU)
    __shared__ __align__(16) __bf16 sq[NB][NPL][128 * RS], sx[NB][NPL][128 * RS];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, half = lane >> 5, col = lane & 31;
    // a 1-D grid of 8 x ncb x ceil(nqb / 8) workgroups: workgroup b runs on XCD b % 8, and every
    // column block of query block qb goes to XCD qb % 8, so a query block is fetched into one L2
    const int ncb = (N + 127) / 128, nqb = (M + 127) / 128;
    const int jb = (int)(blockIdx.x >> 3), cb = jb % ncb, qb = (jb / ncb) * 8 + (int)(blockIdx.x & 7);
    if (qb >= nqb) return; // (uniform over the workgroup, before any barrier)
    const int m0 = qb * 128 + (w >> 1) * 64, n0 = cb * 128 + (w & 1) * 64;
    // Chunks of KC columns of the workgroup's 128 query rows and 128 rows are staged in LDS by
    // coalesced loads (four lanes read a row's 64 contiguous bytes: whole L1 accesses, and each
    // row is fetched once per workgroup, not once per wave that multiplies it); the next chunk's
    // loads are in registers while the waves multiply the current one from LDS and are stored
    // after the products. One buffer (40 KB) and at most 168 registers let three workgroups share
    // a CU, so a 10,000 x 1,024 product's 632 workgroups run in one round on 256 CUs. Staging role of
    // thread t: rows (t >> 2) and (t >> 2) + 64, 16-byte column t & 3 of the chunk; columns at or
    // past dp (the last chunk when dp % 32 == 16) are staged as zeros.
    const int sr = tid >> 2, sc = tid & 3;
    int64_t gq[2], gx[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        gq[u] = (int64_t)min(qb * 128 + sr + 64 * u, M - 1) * dp + 8 * sc;
        gx[u] = (int64_t)min(cb * 128 + sr + 64 * u, N - 1) * dp + 8 * sc;
    }
    bf16x8 rq[NPL][2], rx[NPL][2];
    auto gload = [&](int k0) __attribute__((always_inline)) {
        const bool in = k0 + 8 * sc < dp; // (dp is a multiple of 16: a piece is all in or all out)
        const int kk0 = in ? k0 : 0;
#pragma unroll
        for (int pl = 0; pl < NPL; ++pl)
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const bf16x8 vq = *(const bf16x8 *)((pl == 0 ? Qh : Ql) + gq[u] + kk0);
                const bf16x8 vx = *(const bf16x8 *)((pl == 0 ? Xh : Xl) + gx[u] + kk0);
                rq[pl][u] = in ? vq : bf16x8{};
                rx[pl][u] = in ? vx : bf16x8{};
            }
    };
    auto sstore = [&](int bf) __attribute__((always_inline)) {
#pragma unroll
        for (int pl = 0; pl < NPL; ++pl)
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                *(bf16x8 *)(&sq[bf][pl][(sr + 64 * u) * RS + 8 * sc]) = rq[pl][u];
                *(bf16x8 *)(&sx[bf][pl][(sr + 64 * u) * RS + 8 * sc]) = rx[pl][u];
            }
    };
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x16{};
    const int qr = (w >> 1) * 64 + col, xr = (w & 1) * 64 + col; // the wave's rows in the staged block
    const int nch = (dp + KC - 1) / KC;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int c = 0; c < nch; ++c) {
        const int bf = NB == 1 ? 0 : c & 1;
        if (c + 1 < nch) gload(KC * (c + 1)); // (in flight during the products)
#pragma unroll
        for (int ks = 0; ks < KC / 16; ++ks) {
            bf16x8 a[NPL][2], b[NPL][2];
#pragma unroll
            for (int pl = 0; pl < NPL; ++pl)
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    a[pl][i] = *(const bf16x8 *)(&sq[bf][pl][(qr + 32 * i) * RS + 16 * ks + 8 * half]);
                    b[pl][i] = *(const bf16x8 *)(&sx[bf][pl][(xr + 32 * i) * RS + 16 * ks + 8 * half]);
                }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], b[0][j], acc[i][j], 0, 0, 0);
                    if (PREC == 3) {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], b[NPL - 1][j], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[NPL - 1][i], b[0][j], acc[i][j], 0, 0, 0);
                    }
                }
        }
        if (c + 1 < nch) {
            // two buffers: the other one was last read for chunk c - 1, which every wave finished
            // before the previous barrier; one buffer: wait until every wave is done with chunk c
            if (NB == 1) __syncthreads();
            sstore(NB == 1 ? 0 : bf ^ 1);
            __syncthreads();
        }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int n = n0 + 32 * j + col;
        if (n >= N) continue;
        const float xv = xn[n];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + 32 * i + 8 * (r >> 2) + 4 * half + (r & 3);
                if (m < M) keys[(int64_t)m * N + n] = f32_key(fmaxf(0.f, (qn[m] + xv) - 2.f * acc[i][j][r]));
            }
    }
}

// The first K2c (<= 64) of the `total` (<= 64 E) keys packed in wb, in order, into query q's
// cidx / capx: one bitonic sort of E keys per lane (the K2c smallest end in v[0] of lanes
// 0 .. K2c - 1).
template <int E>
__device__ __forceinline__ void select_sorted_out(const uint64_t *wb, int total, int lane, int64_t q, int K2c,
                                                  int32_t *__restrict__ cidx, float *__restrict__ capx) {
    uint64_t v[E];
#pragma unroll
    for (int u = 0; u < E; ++u) v[u] = lane + 64 * u < total ? wb[lane + 64 * u] : kKey64None;
    wave_sort_u64<E, 2>(v, lane);
    if (lane < K2c) {
        cidx[q * K2c + lane] = v[0] == kKey64None ? -1 : (int32_t)(uint32_t)(v[0] & 0xFFFFFFFFull);
        capx[q * K2c + lane] = key_f32((uint32_t)(v[0] >> 32));
    }
}

// One wave per query: the K2c smallest (key, row) of its key row, ascending, into cidx / capx.
// A lane sorts its keys 16 at a time (rows lane, lane + 64, ...) and keeps its 16 smallest; the
// wave then takes the K2c smallest of the 64 lane lists. capx[K2c - 1] is lowered to the smallest
// key any lane dropped, so it stays a lower bound of every approximation left out.
__global__ __launch_bounds__(256) void l2_matrix_select_kernel(const uint32_t *__restrict__ keys, int N, int64_t M,
                                                               int K2c, int32_t *__restrict__ cidx,
                                                               float *__restrict__ capx) {
    __shared__ uint64_t s_sel[4][256];
    const int lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= M) return;
    const uint32_t *row = keys + q * N;
    // Bounded selection: the K2c-th smallest of the 64 lane minima (one wave sort) bounds the K2c
    // smallest keys from above (the K2c smallest minima are at or below it); the keys at or below
    // it are packed into LDS and sorted, and the first K2c are the result. capx[K2c - 1] is then
    // the K2c-th smallest key, a lower bound of every key left out. More than 256 keys at or below
    // the bound: the chunked sort below.
    if (N <= 16 * 64) { // the row's keys stay in registers (16 per lane) for the three passes
        uint32_t kr[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) kr[j] = lane + 64 * j < N ? row[lane + 64 * j] : kKey32None;
        auto key = [&](int j) { return ((uint64_t)kr[j] << 32) | (uint32_t)(lane + 64 * j); };
        // (the bound on the keys' high words alone: every key whose word is at most the K2c-th
        // smallest lane minimum word is taken, at least K2c keys)
        uint32_t mn = kKey32None;
#pragma unroll
        for (int j = 0; j < 16; ++j) mn = lane + 64 * j < N ? min(mn, kr[j]) : mn;
        uint32_t sm[1] = {mn};
        wave_sort_u32<1>(sm, lane);
        const uint64_t bnd = ((uint64_t)__shfl(sm[0], K2c - 1, 64) << 32) | 0xFFFFFFFFull;
        int held = 0;
#pragma unroll
        for (int j = 0; j < 16; ++j) held += lane + 64 * j < N && key(j) <= bnd ? 1 : 0;
        int off = held;
#pragma unroll
        for (int dd = 1; dd < 64; dd <<= 1) {
            const int t = __shfl_up(off, dd, 64);
            if (lane >= dd) off += t;
        }
        const int total = __shfl(off, 63, 64);
        off -= held;
        if (total <= 256) {
            uint64_t *wb = s_sel[threadIdx.x >> 6];
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (lane + 64 * j < N && key(j) <= bnd) wb[off++] = key(j);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // the packed keys sorted by a network only as wide as their count (usually under 64)
            if (total <= 64)
                select_sorted_out<1>(wb, total, lane, q, K2c, cidx, capx);
            else if (total <= 128)
                select_sorted_out<2>(wb, total, lane, q, K2c, cidx, capx);
            else
                select_sorted_out<4>(wb, total, lane, q, K2c, cidx, capx);
            return;
        }
    } else {
        uint64_t mn = kKey64None;
        for (int c = lane; c < N; c += 64) {
            const uint64_t v = ((uint64_t)row[c] << 32) | (uint32_t)c;
            mn = v < mn ? v : mn;
        }
        uint64_t sm[1] = {mn};
        wave_sort_u64<1, 2>(sm, lane);
        const uint64_t bnd = __shfl(sm[0], K2c - 1, 64);
        int held = 0;
        for (int c = lane; c < N; c += 64) held += (((uint64_t)row[c] << 32) | (uint32_t)c) <= bnd ? 1 : 0;
        int off = held;
#pragma unroll
        for (int dd = 1; dd < 64; dd <<= 1) {
            const int t = __shfl_up(off, dd, 64);
            if (lane >= dd) off += t;
        }
        const int total = __shfl(off, 63, 64);
        off -= held;
        if (total <= 256) {
            uint64_t *wb = s_sel[threadIdx.x >> 6];
            for (int c = lane; c < N; c += 64) {
                const uint64_t v = ((uint64_t)row[c] << 32) | (uint32_t)c;
                if (v <= bnd) wb[off++] = v;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (total <= 64)
                select_sorted_out<1>(wb, total, lane, q, K2c, cidx, capx);
            else if (total <= 128)
                select_sorted_out<2>(wb, total, lane, q, K2c, cidx, capx);
            else
                select_sorted_out<4>(wb, total, lane, q, K2c, cidx, capx);
            return;
        }
    }
    uint64_t L[16], C[16];
    uint64_t dropped = kKey64None;
    for (int c0 = 0; c0 < N; c0 += 16 * 64) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int c = c0 + 64 * j + lane;
            C[j] = c < N ? (((uint64_t)row[c] << 32) | (uint32_t)c) : kKey64None;
        }
        sort16(C);
        if (c0 == 0) {
#pragma unroll
            for (int j = 0; j < 16; ++j) L[j] = C[j];
        } else { // keep the 16 smallest of L and C: bitonic split, then sort the lower half
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const uint64_t a = L[j], b = C[15 - j];
                L[j] = a < b ? a : b;
                const uint64_t hi = a < b ? b : a;
                dropped = hi < dropped ? hi : dropped;
            }
            sort16(L);
        }
    }
    int head = 0;
    for (int r = 0; r < K2c; ++r) {
        uint64_t mine = kKey64None;
#pragma unroll
        for (int j = 0; j < 16; ++j) mine = j == head ? L[j] : mine;
        const uint64_t m = wave_min_u64(mine);
        if (mine == m && m != kKey64None) ++head;
        if (lane == 0) {
            cidx[q * K2c + r] = m == kKey64None ? -1 : (int32_t)(uint32_t)(m & 0xFFFFFFFFull);
            capx[q * K2c + r] = key_f32((uint32_t)(m >> 32));
        }
    }
    const uint64_t dmin = wave_min_u64(dropped);
    if (lane == 0 && dmin != kKey64None) {
        const float dv = key_f32((uint32_t)(dmin >> 32));
        const float last = capx[q * K2c + K2c - 1];
        capx[q * K2c + K2c - 1] = dv < last ? dv : last;
    }
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

// Register-resident query chains in the rerank pay off when few queries leave the machine idle
// and each wave's latency decides (measured: 4,096 queries are faster without).
static bool rerank_reg(int64_t nq) {
    const int64_t t = tune_get("l2.rerank_reg");
    return t == 1 || (t == 2 && nq <= 1024);
}

static int kl_for(int kk) { return kk <= 8 ? 8 : (kk <= 12 ? 16 : 32); }

void l2_mfma_search(const L2Set &xs, const float *X, int64_t xstride, const float *Q, int64_t qstride, int64_t nq,
                    int kk, int32_t *out_idx, float *out_dist, L2Work &wk, int cus, hipStream_t st) {
    const int64_t n = xs.n, d = xs.d;
    const int dp = xs.dp;
    int KL = kl_for(kk);
    int K2 = 2 * KL;
    int K2c = (int)std::min<int64_t>(K2, n);
    const int64_t gx = ceil_div(nq, 128);
    // fallback: a workgroup per query, up to kFallbackMaxGrid workgroups (each then takes every
    // kFallbackMaxGrid-th query), so the grid stays within the launch limits at any query count
    const unsigned fb_grid =
        (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(nq, (int64_t)kFallbackChunk), kFallbackMaxGrid));
    const int64_t nq_pad = gx * 128;
    // queries: split into hi/lo (padded to nq_pad rows of dp; padding rows are zero)
    L2Set &qs = wk.qsplit;
    qs.dp = dp;
    __bf16 *Qh = (__bf16 *)qs.hi.ensure((size_t)nq_pad * dp * 2);
    __bf16 *Ql = (__bf16 *)qs.lo.ensure((size_t)nq_pad * dp * 2);
    float *qn = (float *)qs.norm.ensure((size_t)nq_pad * 4);
    hipLaunchKernelGGL(l2_split_kernel, dim3((unsigned)ceil_div(nq_pad, 4)), dim3(256), 0, st, Q, nq, nq_pad, d,
                       qstride, dp, Qh, Ql, qn);
    const int prec_m = (int)tune_get("l2.matrix_terms");
    if (n <= 4096 && dp > 64 && tune_get("l2.matrix") != 0) {
        // small database of wide rows: key matrix per query chunk (<= 256 MB), then selection
        int32_t *cidx = (int32_t *)wk.cidx.ensure((size_t)nq * K2c * 4);
        float *capx = (float *)wk.capx.ensure((size_t)nq * K2c * 4);
        int32_t *flags = (int32_t *)wk.flags.ensure((size_t)nq * 4);
        const int64_t chunk = std::max<int64_t>(128, std::min<int64_t>(nq_pad, ((int64_t)1 << 26) / n / 128 * 128));
        uint32_t *keys = (uint32_t *)wk.part.ensure((size_t)chunk * n * 4);
        for (int64_t c0 = 0; c0 < nq; c0 += chunk) {
            const int64_t m = std::min<int64_t>(chunk, nq - c0);
            {
                StageScope sc(st, STAGE_SCAN);
                dim3 grid((unsigned)(8 * ceil_div(n, (int64_t)128) * ceil_div(ceil_div(m, (int64_t)128), (int64_t)8)));
                if (prec_m == 1)
                    hipLaunchKernelGGL((l2_matrix_keys_kernel<1>), grid, dim3(256), 0, st, Qh + c0 * dp, Ql + c0 * dp,
                                       qn + c0, (int)m, xs.hi.as<__bf16>(), xs.lo.as<__bf16>(), xs.norm.as<float>(),
                                       (int)n, dp, keys);
                else
                    hipLaunchKernelGGL((l2_matrix_keys_kernel<3>), grid, dim3(256), 0, st, Qh + c0 * dp, Ql + c0 * dp,
                                       qn + c0, (int)m, xs.hi.as<__bf16>(), xs.lo.as<__bf16>(), xs.norm.as<float>(),
                                       (int)n, dp, keys);
            }
            {
                StageScope sc(st, STAGE_SELECT);
                dim3 g((unsigned)ceil_div(m, (int64_t)4));
                hipLaunchKernelGGL(l2_matrix_select_kernel, g, dim3(256), 0, st, keys, (int)n, m, K2c, cidx + c0 * K2c,
                                   capx + c0 * K2c);
            }
            PN_HIP(hipGetLastError());
        }
        {
            StageScope sc(st, STAGE_RERANK);
            const float eps_coef = (prec_m == 1 ? 0.0078125f : 6.103515625e-05f) + (float)dp * 4.76837158203125e-07f;
            hipLaunchKernelGGL(rerank_reg(nq) ? l2_rerank_kernel<true> : l2_rerank_kernel<false>, dim3((unsigned)ceil_div(nq, (int64_t)4)), dim3(256), 0, st, X, xstride,
                               d, Q, qstride, qn, xs.norm_max.as<unsigned int>(), cidx, capx, K2c,
                               (const uint64_t *)nullptr, (int64_t)0, KL, nq_pad, nq, kk, n, eps_coef, out_idx,
                               out_dist, flags);
            PN_HIP(hipGetLastError());
        }
        {
            StageScope sc(st, STAGE_OTHER);
            if (kk <= 8)
                hipLaunchKernelGGL((l2_fallback_kernel<8>), dim3(fb_grid), dim3(256), 0, st, X, xstride, n, d, Q,
                                   qstride, flags, kk, out_idx, out_dist, nq);
            else if (kk <= 16)
                hipLaunchKernelGGL((l2_fallback_kernel<16>), dim3(fb_grid), dim3(256), 0, st, X, xstride, n, d, Q,
                                   qstride, flags, kk, out_idx, out_dist, nq);
            else
                hipLaunchKernelGGL((l2_fallback_kernel<32>), dim3(fb_grid), dim3(256), 0, st, X, xstride, n, d, Q,
                                   qstride, flags, kk, out_idx, out_dist, nq);
            PN_HIP(hipGetLastError());
        }
        return;
    }
    // slices: whole multiples of the CU count, at least 128 rows each
    const int64_t wpc = std::max<int64_t>(1, tune_get("l2.wg_per_cu"));
    int64_t S = std::max<int64_t>(1, wpc * cus / gx);
    S = std::min<int64_t>(S, std::max<int64_t>(1, n / 128));
    const int64_t rps = ceil_div(ceil_div(n, S), 32) * 32; // slices start on 32-row boundaries
    S = ceil_div(n, rps);
    const int64_t nlists = S;
    // with at least 4 slices, lists of 4 give K2c = 8 candidates, enough for k <= 6 (fewer keys per
    // lane to keep sorted)
    if (kk <= 6 && S >= 4 && tune_get("l2.kl4") != 0) {
        KL = 4;
        K2 = 8;
        K2c = (int)std::min<int64_t>(K2, n);
    }
    // per-slice lists, query-major: entry j of slice s's list of query q at (q S + s) KL + j
    uint64_t *part = (uint64_t *)wk.part.ensure((size_t)nlists * KL * nq_pad * 8);
    const int64_t qsl = nlists * KL;
    const int prec = (int)tune_get("l2.mfma_terms");
    {
        StageScope sc(st, STAGE_SCAN);
        // slice-major when the query blocks fit the grid's y dimension (see the kernels)
        const bool sm = gx <= 65535 && tune_get("l2.slice_major") != 0;
        dim3 grid(sm ? (unsigned)S : (unsigned)gx, sm ? (unsigned)gx : (unsigned)S);
        const bool wide = tune_get("l2.wide") != 0; // register-resident queries for dp = 128, 192, 256
        auto launch = [&](auto klc, auto pc) {
            constexpr int KLV = decltype(klc)::value, PV = decltype(pc)::value;
            const __bf16 *xh = xs.hi.as<__bf16>(), *xl = xs.lo.as<__bf16>();
            const float *xnr = xs.norm.as<float>();
            switch (dp) {
            case 16:
                hipLaunchKernelGGL((l2_mfma_scan_small_kernel<KLV, PV, 1>), grid, dim3(256), 0, st, xh, xl, xnr, n, Qh,
                                   Ql, qn, rps, part, qsl, sm);
                break;
            case 32:
                hipLaunchKernelGGL((l2_mfma_scan_small_kernel<KLV, PV, 2>), grid, dim3(256), 0, st, xh, xl, xnr, n, Qh,
                                   Ql, qn, rps, part, qsl, sm);
                break;
            case 48:
                hipLaunchKernelGGL((l2_mfma_scan_small_kernel<KLV, PV, 3>), grid, dim3(256), 0, st, xh, xl, xnr, n, Qh,
                                   Ql, qn, rps, part, qsl, sm);
                break;
            case 64:
                hipLaunchKernelGGL((l2_mfma_scan_small_kernel<KLV, PV, 4>), grid, dim3(256), 0, st, xh, xl, xnr, n, Qh,
                                   Ql, qn, rps, part, qsl, sm);
                break;
            case 128:
            case 192:
            case 256:
                if (wide) {
                    if (dp == 128)
                        hipLaunchKernelGGL((l2_mfma_scan_wide_kernel<KLV, PV, 2>), grid, dim3(256), 0, st, xh, xl, xnr,
                                           n, Qh, Ql, qn, rps, part, qsl, sm);
                    else if (dp == 192)
                        hipLaunchKernelGGL((l2_mfma_scan_wide_kernel<KLV, PV, 3>), grid, dim3(256), 0, st, xh, xl, xnr,
                                           n, Qh, Ql, qn, rps, part, qsl, sm);
                    else
                        hipLaunchKernelGGL((l2_mfma_scan_wide_kernel<KLV, PV, 4>), grid, dim3(256), 0, st, xh, xl, xnr,
                                           n, Qh, Ql, qn, rps, part, qsl, sm);
                    break;
                }
                [[fallthrough]];
            default:
                hipLaunchKernelGGL((l2_mfma_scan_kernel<KLV, PV>), grid, dim3(256), 0, st, xh, xl, xnr, n, dp, Qh, Ql, qn,
                                   rps, part, qsl, sm);
            }
        };
        using I4 = std::integral_constant<int, 4>;
        using I8 = std::integral_constant<int, 8>;
        using I16 = std::integral_constant<int, 16>;
        using I32 = std::integral_constant<int, 32>;
        using P1 = std::integral_constant<int, 1>;
        using P3 = std::integral_constant<int, 3>;
        if (prec == 1) {
            if (KL == 4) launch(I4{}, P1{}); else if (KL == 8) launch(I8{}, P1{});
            else if (KL == 16) launch(I16{}, P1{}); else launch(I32{}, P1{});
        } else {
            if (KL == 4) launch(I4{}, P3{}); else if (KL == 8) launch(I8{}, P3{});
            else if (KL == 16) launch(I16{}, P3{}); else launch(I32{}, P3{});
        }
        PN_HIP(hipGetLastError());
    }
    int32_t *cidx = (int32_t *)wk.cidx.ensure((size_t)nq * K2c * 4);
    float *capx = (float *)wk.capx.ensure((size_t)nq * K2c * 4);
    int32_t *flags = (int32_t *)wk.flags.ensure((size_t)nq * 4);
    const float eps_coef = (prec == 1 ? 0.0078125f : 6.103515625e-05f) + (float)dp * 4.76837158203125e-07f;
    // a lane that merges one slice holds at most KL keys; with more slices per lane it needs K2c
    const int cap = nlists <= 64 ? KL : K2;
    const int64_t lt = tune_get("merge.lean");
    const bool lean = K2c <= 64 && (lt >= 2 || (lt == 1 && cap >= 16)); // as merge_lists_wave
    if (tune_get("l2.fuse_rerank") != 0 && (lean || cap == 4 || cap == 8)) {
        StageScope sc(st, STAGE_SELECT);
        const bool reg = rerank_reg(nq);
        auto go = [&](auto kc, auto rc) {
            hipLaunchKernelGGL((l2_merge_rerank_kernel<decltype(kc)::value, decltype(rc)::value>),
                               dim3(xcd_grid(ceil_div(nq, (int64_t)4))), dim3(256), 0, st, part, nlists, KL, qsl, K2c,
                               lt == 2, X, xstride, d, Q, qstride, qn, xs.norm_max.as<unsigned int>(), nq, kk, n,
                               eps_coef, out_idx, out_dist, flags);
        };
        using R0 = std::false_type;
        using R1 = std::true_type;
        const int kc = lean ? 0 : cap;
        if (kc == 0) reg ? go(std::integral_constant<int, 0>{}, R1{}) : go(std::integral_constant<int, 0>{}, R0{});
        else if (kc == 4) reg ? go(std::integral_constant<int, 4>{}, R1{}) : go(std::integral_constant<int, 4>{}, R0{});
        else reg ? go(std::integral_constant<int, 8>{}, R1{}) : go(std::integral_constant<int, 8>{}, R0{});
        PN_HIP(hipGetLastError());
    } else {
        {
            StageScope sc(st, STAGE_SELECT);
            merge_lists_wave(part, nlists, KL, cap, nq, qsl, 1, K2c, OUT_F32, cidx, capx, st);
        }
        StageScope sc(st, STAGE_RERANK);
        hipLaunchKernelGGL(rerank_reg(nq) ? l2_rerank_kernel<true> : l2_rerank_kernel<false>, dim3((unsigned)ceil_div(nq, (int64_t)4)), dim3(256), 0, st, X, xstride, d, Q,
                           qstride, qn, xs.norm_max.as<unsigned int>(), cidx, capx, K2c, part, nlists, KL, qsl, nq,
                           kk, n, eps_coef, out_idx, out_dist, flags);
        PN_HIP(hipGetLastError());
    }
    {
        StageScope sc(st, STAGE_OTHER);
#define PN_L2FB(KV)                                                                                              \
    hipLaunchKernelGGL((l2_fallback_kernel<KV>), dim3(fb_grid), dim3(256), 0, st, X, xstride, n, d, Q, qstride, \
                       flags, kk, out_idx, out_dist, nq)
        if (kk <= 8) PN_L2FB(8); else if (kk <= 16) PN_L2FB(16); else PN_L2FB(32);
#undef PN_L2FB
        PN_HIP(hipGetLastError());
    }
}

} // namespace pn
