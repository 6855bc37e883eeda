// pn_probe: a learned probe predictor for the IVF-L2 search -- an MLP d -> hidden (ReLU) -> nlist
// in bf16 with fp32 accumulation, trained on the device by the kernels of this file, and an
// adaptive probe selection from its softmax.
//
// No reference counterpart: the reference probes the n_probe nearest centroids of every query
// (forest.py:76-80). Here a query's probes are the clusters the network ranks highest, as many as
// it takes for their predicted probability mass to reach a target (between np_min and np_max).
//
// Training target: for a query with exact top-L neighbours whose clusters are c_1..c_L, the
// target distribution puts mass 1/L on each c_j (repeats add up); the loss is the cross-entropy
// of softmax(logits) against it, averaged over the batch.
//
// Layout: weights are kept in fp32 (masters, Adam moments) and as bf16 copies in "k-contiguous"
// form for the products: W1 [hidden][d], W2 [nlist][hidden], and W2T [hidden][nlist] for the
// backward product. Every product is C[m][n] = sum_k A[m][k] B[n][k] (both operands k-contiguous)
// on v_mfma_f32_32x32x16_bf16; the other orientations are made by transposing bf16 copies.
#include <cmath>
#include <memory>
#include <random>
#include <vector>

#include "common.hpp"
#include "device_util.hpp"
#include "l2_mfma.hpp"

namespace pn {

enum { EPI_F32 = 0, EPI_BIAS_RELU_BF16 = 1, EPI_BIAS_F32 = 2, EPI_RELU_MASK_BF16 = 3 };

struct EpiArgs {
    float *out_f32;            // EPI_F32, EPI_BIAS_F32
    __bf16 *out_bf16;          // EPI_BIAS_RELU_BF16, EPI_RELU_MASK_BF16
    const float *bias;         // per column n
    const __bf16 *mask;        // EPI_RELU_MASK_BF16: values > 0 keep the gradient
    int64_t ldc;
};

// C[M x N] = A[M x K] . B[N x K]^T, bf16 operands (row strides lda, ldb), K % 16 == 0. A
// workgroup computes a 128 x 128 block; wave w a 64 x 64 quarter as 2 x 2 MFMA tiles. Rows past
// M or N are clamped on load and masked on store.
template <int EPI>
__global__ __launch_bounds__(256) void gemm_nt_kernel(const __bf16 *__restrict__ A, int64_t lda,
                                                      const __bf16 *__restrict__ B, int64_t ldb, int M, int N, int K,
                                                      EpiArgs ep) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, half = lane >> 5, col = lane & 31;
    const int m0 = blockIdx.y * 128 + (w >> 1) * 64, n0 = blockIdx.x * 128 + (w & 1) * 64;
    const __bf16 *ap[2], *bp[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        ap[i] = A + (int64_t)min(m0 + 32 * i + col, M - 1) * lda + 8 * half;
        bp[i] = B + (int64_t)min(n0 + 32 * i + col, N - 1) * ldb + 8 * half;
    }
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x16{};
    int k0 = 0;
    for (; k0 + 64 <= K; k0 += 64) {
        bf16x8 a[2][4], b[2][4];
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                a[i][s] = *(const bf16x8 *)(ap[i] + k0 + 16 * s);
                b[i][s] = *(const bf16x8 *)(bp[i] + k0 + 16 * s);
            }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][s], b[j][s], acc[i][j], 0, 0, 0);
    }
    for (; k0 < K; k0 += 16) {
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            a[i] = *(const bf16x8 *)(ap[i] + k0);
            b[i] = *(const bf16x8 *)(bp[i] + k0);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    // D layout: lane (col, half), register r -> row 8(r/4) + 4 half + r%4, column col
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int n = n0 + 32 * j + col;
        if (n >= N) continue;
        const float bias = (EPI == EPI_BIAS_RELU_BF16 || EPI == EPI_BIAS_F32) ? ep.bias[n] : 0.f;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + 32 * i + 8 * (r >> 2) + 4 * half + (r & 3);
                if (m >= M) continue;
                const int64_t o = (int64_t)m * ep.ldc + n;
                const float v = acc[i][j][r] + bias;
                if (EPI == EPI_F32 || EPI == EPI_BIAS_F32) ep.out_f32[o] = v;
                if (EPI == EPI_BIAS_RELU_BF16) ep.out_bf16[o] = (__bf16)fmaxf(v, 0.f);
                if (EPI == EPI_RELU_MASK_BF16) ep.out_bf16[o] = (float)ep.mask[o] > 0.f ? (__bf16)v : (__bf16)0.f;
            }
    }
}

template <int EPI>
static void gemm_nt(const __bf16 *A, int64_t lda, const __bf16 *B, int64_t ldb, int M, int N, int K, EpiArgs ep,
                    hipStream_t st) {
    PN_REQUIRE(K % 16 == 0, PN_EINVAL, "product depth must be a multiple of 16");
    dim3 grid((unsigned)ceil_div(N, 128), (unsigned)ceil_div(M, 128));
    hipLaunchKernelGGL((gemm_nt_kernel<EPI>), grid, dim3(256), 0, st, A, lda, B, ldb, M, N, K, ep);
    PN_HIP(hipGetLastError());
}

// out[c][r] = in[r][c] (bf16, rows x cols), 64 x 64 tiles through LDS.
__global__ __launch_bounds__(256) void transpose_bf16_kernel(const __bf16 *__restrict__ in, int rows, int cols,
                                                             __bf16 *__restrict__ out) {
    __shared__ __bf16 t[64][65];
    const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int rr = e / 64, cc = e % 64;
        if (r0 + rr < rows && c0 + cc < cols) t[rr][cc] = in[(int64_t)(r0 + rr) * cols + c0 + cc];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int cc = e / 64, rr = e % 64;
        if (r0 + rr < rows && c0 + cc < cols) out[(int64_t)(c0 + cc) * rows + r0 + rr] = t[rr][cc];
    }
}

static void transpose_bf16(const __bf16 *in, int rows, int cols, __bf16 *out, hipStream_t st) {
    dim3 grid((unsigned)ceil_div(cols, 64), (unsigned)ceil_div(rows, 64));
    hipLaunchKernelGGL(transpose_bf16_kernel, grid, dim3(256), 0, st, in, rows, cols, out);
}

__global__ void to_bf16_kernel(const float *__restrict__ x, int64_t n, __bf16 *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (__bf16)x[i];
}

__device__ __forceinline__ float block_sum(float v, float *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

__device__ __forceinline__ float block_max(float v, float *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// One workgroup per row: softmax of logits[row] (ncls), cross-entropy against the target of the
// row's L labels, gradient (p - t) * scale as bf16, loss * scale added to *loss.
__global__ __launch_bounds__(256) void softmax_xent_kernel(const float *__restrict__ logits, int ncls,
                                                           const int32_t *__restrict__ labels, int L, float scale,
                                                           __bf16 *__restrict__ grad, float *__restrict__ loss) {
    __shared__ float red[4];
    const int64_t row = blockIdx.x;
    const float *x = logits + row * ncls;
    const int32_t *lab = labels + row * L;
    float mx = -INFINITY;
    for (int c = threadIdx.x; c < ncls; c += 256) mx = fmaxf(mx, x[c]);
    mx = block_max(mx, red);
    float se = 0.f;
    for (int c = threadIdx.x; c < ncls; c += 256) se += expf(x[c] - mx);
    se = block_sum(se, red);
    const float lse = mx + logf(se);
    const float inv_l = 1.f / (float)L;
    for (int c = threadIdx.x; c < ncls; c += 256) {
        float t = 0.f;
        for (int j = 0; j < L; ++j) t += lab[j] == c ? inv_l : 0.f;
        grad[row * ncls + c] = (__bf16)((expf(x[c] - lse) - t) * scale);
    }
    if (threadIdx.x == 0) {
        float l = 0.f;
        for (int j = 0; j < L; ++j)
            if (lab[j] >= 0 && lab[j] < ncls) l += inv_l * (lse - x[lab[j]]);
        atomicAdd(loss, l * scale);
    }
}

// out[c] = sum over rows of in[r][c] (bf16 in, fp32 out), one thread per column.
__global__ void colsum_bf16_kernel(const __bf16 *__restrict__ in, int rows, int cols, float *__restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; ++r) s += (float)in[(int64_t)r * cols + c];
    out[c] = s;
}

// Adam (bias-corrected) on fp32 masters; refreshes the bf16 copy.
__global__ void adam_kernel(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1,
                            float c2, __bf16 *__restrict__ wb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = (__bf16)wi;
}

// One workgroup per query: softmax of its logits, then the clusters with the highest predicted
// probability: the fewest whose mass reaches `mass`, at least np_min and at most np_max (a
// threshold on the probability found by bisection; clusters tied at the threshold fill the
// slots left below np_max in no fixed order). Writes probes[q][0..count) and counts[q].
__global__ __launch_bounds__(256) void select_probes_kernel(const float *__restrict__ logits, int ncls, float mass,
                                                            int np_min, int np_max, int32_t *__restrict__ probes,
                                                            int32_t *__restrict__ counts) {
    __shared__ float red[4];
    __shared__ int s_n;
    const int64_t row = blockIdx.x;
    const float *x = logits + row * ncls;
    float mx = -INFINITY;
    for (int c = threadIdx.x; c < ncls; c += 256) mx = fmaxf(mx, x[c]);
    mx = block_max(mx, red);
    float se = 0.f;
    for (int c = threadIdx.x; c < ncls; c += 256) se += expf(x[c] - mx);
    se = block_sum(se, red);
    const float inv = 1.f / se;
    // bisection on t in [0, 1]: mass(t) = sum of p >= t, cnt(t) = #{p >= t}; both fall as t rises
    float lo = 0.f, hi = 1.f; // invariant: the answer t* lies in [lo, hi]
    for (int it = 0; it < 30; ++it) {
        const float t = 0.5f * (lo + hi);
        float ms = 0.f, cn = 0.f;
        for (int c = threadIdx.x; c < ncls; c += 256) {
            const float p = expf(x[c] - mx) * inv;
            if (p >= t) {
                ms += p;
                cn += 1.f;
            }
        }
        ms = block_sum(ms, red);
        cn = block_sum(cn, red);
        // t is acceptable (no lower than needed) when its set is within bounds and reaches the mass
        const bool enough = (ms >= mass && cn >= (float)np_min) || cn >= (float)np_max;
        if (enough) lo = t; else hi = t;
    }
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    for (int c = threadIdx.x; c < ncls; c += 256) {
        const float p = expf(x[c] - mx) * inv;
        if (p >= lo) {
            const int slot = atomicAdd(&s_n, 1);
            if (slot < np_max) probes[row * np_max + slot] = c;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) counts[row] = min(s_n, np_max);
}

// The same selection with one wave per query for ncls <= 64 J: the probabilities stay in registers
// (J per lane, cluster lane + 64 j) for the whole bisection, whose mass and count sums are wave
// sums by DPP moves, so a step costs three operations per cluster and no barrier. The chosen
// clusters are written in ascending order (ballot compaction).
template <int J>
__global__ __launch_bounds__(256) void select_probes_wave_kernel(const float *__restrict__ logits, int ncls, float mass,
                                                                 int np_min, int np_max, int64_t B,
                                                                 int32_t *__restrict__ probes,
                                                                 int32_t *__restrict__ counts) {
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= B) return;
    const float *x = logits + row * ncls;
    float p[J];
    float mx = -INFINITY;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int c = lane + 64 * j;
        p[j] = c < ncls ? x[c] : -INFINITY;
        mx = fmaxf(mx, p[j]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    float se = 0.f;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        p[j] = expf(p[j] - mx); // (absent clusters: exp(-inf) = 0, never selected: t > 0 below)
        se += p[j];
    }
    se = wave_sum_dpp(se);
    const float inv = 1.f / se;
    float pmax = 0.f;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        p[j] *= inv;
        pmax = fmaxf(pmax, p[j]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) pmax = fmaxf(pmax, __shfl_xor(pmax, off, 64));
    // bisection on t in [0, pmax] as in select_probes_kernel (t = pmax always selects at least one
    // cluster, so the answer lies below it or is it)
    float lo = 0.f, hi = pmax;
    for (int it = 0; it < 30; ++it) {
        const float t = 0.5f * (lo + hi);
        float ms = 0.f, cn = 0.f;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const bool in = p[j] >= t;
            ms += in ? p[j] : 0.f;
            cn += in ? 1.f : 0.f;
        }
        ms = wave_sum_dpp(ms);
        cn = wave_sum_dpp(cn);
        const bool enough = (ms >= mass && cn >= (float)np_min) || cn >= (float)np_max;
        if (enough) lo = t; else hi = t;
    }
    int base = 0;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const bool sel = p[j] >= lo && lane + 64 * j < ncls;
        const uint64_t bal = __ballot(sel);
        const int pos = base + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
        if (sel && pos < np_max) probes[row * np_max + pos] = lane + 64 * j;
        base += __popcll(bal);
    }
    if (lane == 0) counts[row] = min(base, np_max);
}

} // namespace pn

struct pn_probe {
    int device = 0;
    int d = 0, hid = 0, ncls = 0;
    int64_t step = 0;
    // fp32 masters and Adam moments, packed [W1 | b1 | W2 | b2]
    pn::DevBuf w, g, m, v;
    pn::DevBuf W1h, W2h, W2Th; // bf16 copies
    // per batch
    pn::DevBuf Xh, Xt, Hh, Ht, logits, dL, dLt, dH, dHt, loss;
    int64_t cap = 0;
    int64_t n1() const { return (int64_t)hid * d; }
    int64_t n2() const { return (int64_t)ncls * hid; }
    int64_t total() const { return n1() + hid + n2() + ncls; }
    float *W1() const { return w.as<float>(); }
    float *b1() const { return w.as<float>() + n1(); }
    float *W2() const { return w.as<float>() + n1() + hid; }
    float *b2() const { return w.as<float>() + n1() + hid + n2(); }
};

namespace pn {

static void probe_reserve(pn_probe *p, int64_t B) {
    if (B <= p->cap) return;
    p->Xh.ensure((size_t)B * p->d * 2);
    p->Xt.ensure((size_t)B * p->d * 2);
    p->Hh.ensure((size_t)B * p->hid * 2);
    p->Ht.ensure((size_t)B * p->hid * 2);
    p->logits.ensure((size_t)B * p->ncls * 4);
    p->dL.ensure((size_t)B * p->ncls * 2);
    p->dLt.ensure((size_t)B * p->ncls * 2);
    p->dH.ensure((size_t)B * p->hid * 2);
    p->dHt.ensure((size_t)B * p->hid * 2);
    p->cap = B;
}

static void refresh_bf16(pn_probe *p, hipStream_t st) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)ceil_div(p->n1(), (int64_t)256)), dim3(256), 0, st, p->W1(),
                       p->n1(), p->W1h.as<__bf16>());
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)ceil_div(p->n2(), (int64_t)256)), dim3(256), 0, st, p->W2(),
                       p->n2(), p->W2h.as<__bf16>());
    transpose_bf16(p->W2h.as<__bf16>(), p->ncls, p->hid, p->W2Th.as<__bf16>(), st);
}

// logits [B x ncls] of device queries x (fp32 B x d)
static void probe_forward(pn_probe *p, const float *x, int64_t B, hipStream_t st) {
    probe_reserve(p, B);
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)ceil_div(B * p->d, (int64_t)256)), dim3(256), 0, st, x,
                       B * p->d, p->Xh.as<__bf16>());
    EpiArgs e1{nullptr, p->Hh.as<__bf16>(), p->b1(), nullptr, p->hid};
    gemm_nt<EPI_BIAS_RELU_BF16>(p->Xh.as<__bf16>(), p->d, p->W1h.as<__bf16>(), p->d, (int)B, p->hid, p->d, e1, st);
    EpiArgs e2{p->logits.as<float>(), nullptr, p->b2(), nullptr, p->ncls};
    gemm_nt<EPI_BIAS_F32>(p->Hh.as<__bf16>(), p->hid, p->W2h.as<__bf16>(), p->hid, (int)B, p->ncls, p->hid, e2, st);
}

// forward + backward: loss (device float, averaged over the batch) and gradients into p->g
static void probe_backward(pn_probe *p, const float *x, const int32_t *labels, int L, int64_t B, hipStream_t st) {
    PN_REQUIRE(B % 16 == 0, PN_EINVAL, "training batch must be a multiple of 16");
    probe_forward(p, x, B, st);
    float *loss = (float *)p->loss.ensure(4);
    PN_HIP(hipMemsetAsync(loss, 0, 4, st));
    hipLaunchKernelGGL(softmax_xent_kernel, dim3((unsigned)B), dim3(256), 0, st, p->logits.as<float>(), p->ncls,
                       labels, L, 1.f / (float)B, p->dL.as<__bf16>(), loss);
    float *gW1 = p->g.as<float>(), *gb1 = gW1 + p->n1(), *gW2 = gb1 + p->hid, *gb2 = gW2 + p->n2();
    // dH = (dL . W2) masked by ReLU'(H)
    EpiArgs eh{nullptr, p->dH.as<__bf16>(), nullptr, p->Hh.as<__bf16>(), p->hid};
    gemm_nt<EPI_RELU_MASK_BF16>(p->dL.as<__bf16>(), p->ncls, p->W2Th.as<__bf16>(), p->ncls, (int)B, p->hid, p->ncls,
                                eh, st);
    // dW2 = dL^T . H, dW1 = dH^T . X (products over the batch)
    transpose_bf16(p->dL.as<__bf16>(), (int)B, p->ncls, p->dLt.as<__bf16>(), st);
    transpose_bf16(p->Hh.as<__bf16>(), (int)B, p->hid, p->Ht.as<__bf16>(), st);
    transpose_bf16(p->dH.as<__bf16>(), (int)B, p->hid, p->dHt.as<__bf16>(), st);
    transpose_bf16(p->Xh.as<__bf16>(), (int)B, p->d, p->Xt.as<__bf16>(), st);
    EpiArgs e2{gW2, nullptr, nullptr, nullptr, p->hid};
    gemm_nt<EPI_F32>(p->dLt.as<__bf16>(), B, p->Ht.as<__bf16>(), B, p->ncls, p->hid, (int)B, e2, st);
    EpiArgs e1{gW1, nullptr, nullptr, nullptr, p->d};
    gemm_nt<EPI_F32>(p->dHt.as<__bf16>(), B, p->Xt.as<__bf16>(), B, p->hid, p->d, (int)B, e1, st);
    hipLaunchKernelGGL(colsum_bf16_kernel, dim3((unsigned)ceil_div(p->ncls, 256)), dim3(256), 0, st,
                       p->dL.as<__bf16>(), (int)B, p->ncls, gb2);
    hipLaunchKernelGGL(colsum_bf16_kernel, dim3((unsigned)ceil_div(p->hid, 256)), dim3(256), 0, st,
                       p->dH.as<__bf16>(), (int)B, p->hid, gb1);
    PN_HIP(hipGetLastError());
}

} // namespace pn

using namespace pn;

extern "C" pn_status pn_probe_create(int32_t d, int32_t hidden, int32_t nlist, uint32_t seed, pn_probe **out) {
    return guard([&] {
        PN_REQUIRE(out && d > 0 && hidden > 0 && nlist > 0, PN_EINVAL, "bad shape");
        PN_REQUIRE(d % 16 == 0 && hidden % 16 == 0, PN_EINVAL, "d and hidden must be multiples of 16");
        auto *p = new pn_probe();
        std::unique_ptr<pn_probe> hold(p);
        p->device = current_device();
        p->d = d;
        p->hid = hidden;
        p->ncls = nlist;
        const int64_t T = p->total();
        std::vector<float> h(T, 0.f);
        std::mt19937 rng(seed);
        std::normal_distribution<float> nd(0.f, 1.f);
        const float s1 = std::sqrt(2.f / (float)d), s2 = std::sqrt(1.f / (float)hidden);
        for (int64_t i = 0; i < p->n1(); ++i) h[i] = nd(rng) * s1;
        for (int64_t i = 0; i < p->n2(); ++i) h[p->n1() + hidden + i] = nd(rng) * s2;
        hipStream_t st = lib_stream();
        PN_HIP(hipMemcpyAsync(p->w.ensure(T * 4), h.data(), T * 4, hipMemcpyHostToDevice, st));
        PN_HIP(hipMemsetAsync(p->g.ensure(T * 4), 0, T * 4, st));
        PN_HIP(hipMemsetAsync(p->m.ensure(T * 4), 0, T * 4, st));
        PN_HIP(hipMemsetAsync(p->v.ensure(T * 4), 0, T * 4, st));
        p->W1h.ensure(p->n1() * 2);
        p->W2h.ensure(p->n2() * 2);
        p->W2Th.ensure(p->n2() * 2);
        refresh_bf16(p, st);
        PN_HIP(hipStreamSynchronize(st));
        *out = hold.release();
    });
}

extern "C" void pn_probe_destroy(pn_probe *p) {
    if (!p) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(p->device);
    delete p;
    if (prev >= 0) (void)hipSetDevice(prev);
}

extern "C" pn_status pn_probe_get_params(const pn_probe *p, float *params) {
    return guard([&] {
        PN_REQUIRE(p && params, PN_EINVAL, "null pointer");
        DeviceScope ds(p->device);
        PN_HIP(hipMemcpy(params, p->w.p, p->total() * 4, hipMemcpyDeviceToHost));
    });
}

extern "C" pn_status pn_probe_set_params(pn_probe *p, const float *params) {
    return guard([&] {
        PN_REQUIRE(p && params, PN_EINVAL, "null pointer");
        DeviceScope ds(p->device);
        hipStream_t st = lib_stream();
        PN_HIP(hipMemcpyAsync(p->w.p, params, p->total() * 4, hipMemcpyHostToDevice, st));
        refresh_bf16(p, st);
        PN_HIP(hipStreamSynchronize(st));
    });
}

extern "C" pn_status pn_probe_grads_device(pn_probe *p, const float *x, const int32_t *labels, int32_t L, int64_t B,
                                           float *loss_out, float *grads_out, void *stream) {
    return guard([&] {
        PN_REQUIRE(p && x && labels && L > 0 && B > 0, PN_EINVAL, "bad arguments");
        DeviceScope ds(p->device);
        hipStream_t st = (hipStream_t)stream;
        probe_backward(p, x, labels, L, B, st);
        if (loss_out) PN_HIP(hipMemcpyAsync(loss_out, p->loss.p, 4, hipMemcpyDeviceToDevice, st));
        if (grads_out) PN_HIP(hipMemcpyAsync(grads_out, p->g.p, p->total() * 4, hipMemcpyDeviceToDevice, st));
    });
}

extern "C" pn_status pn_probe_train_step_device(pn_probe *p, const float *x, const int32_t *labels, int32_t L,
                                                int64_t B, float lr, float *loss_out, void *stream) {
    return guard([&] {
        PN_REQUIRE(p && x && labels && L > 0 && B > 0, PN_EINVAL, "bad arguments");
        DeviceScope ds(p->device);
        hipStream_t st = (hipStream_t)stream;
        probe_backward(p, x, labels, L, B, st);
        p->step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)p->step), c2 = 1.f - std::pow(b2, (float)p->step);
        const int64_t T = p->total();
        hipLaunchKernelGGL(adam_kernel, dim3((unsigned)ceil_div(T, (int64_t)256)), dim3(256), 0, st, p->w.as<float>(),
                           p->g.as<float>(), p->m.as<float>(), p->v.as<float>(), T, lr, b1, b2, 1e-8f, c1, c2,
                           (__bf16 *)nullptr);
        refresh_bf16(p, st);
        if (loss_out) PN_HIP(hipMemcpyAsync(loss_out, p->loss.p, 4, hipMemcpyDeviceToDevice, st));
        PN_HIP(hipGetLastError());
    });
}

extern "C" pn_status pn_probe_forward_device(pn_probe *p, const float *x, int64_t B, float *logits_out, void *stream) {
    return guard([&] {
        PN_REQUIRE(p && x && logits_out && B > 0, PN_EINVAL, "bad arguments");
        DeviceScope ds(p->device);
        hipStream_t st = (hipStream_t)stream;
        probe_forward(p, x, B, st);
        PN_HIP(hipMemcpyAsync(logits_out, p->logits.p, (size_t)B * p->ncls * 4, hipMemcpyDeviceToDevice, st));
    });
}

extern "C" pn_status pn_probe_select_device(pn_probe *p, const float *x, int64_t B, float mass, int32_t np_min,
                                            int32_t np_max, int32_t *probes, int32_t *counts, void *stream) {
    return guard([&] {
        PN_REQUIRE(p && x && probes && counts && B > 0, PN_EINVAL, "bad arguments");
        PN_REQUIRE(np_min >= 1 && np_max >= np_min && np_max <= p->ncls, PN_EINVAL, "need 1 <= np_min <= np_max <= nlist");
        DeviceScope ds(p->device);
        hipStream_t st = (hipStream_t)stream;
        probe_forward(p, x, B, st);
        const dim3 gw((unsigned)ceil_div(B, (int64_t)4));
        if (p->ncls <= 1024)
            hipLaunchKernelGGL(select_probes_wave_kernel<16>, gw, dim3(256), 0, st, p->logits.as<float>(), p->ncls, mass,
                               np_min, np_max, B, probes, counts);
        else if (p->ncls <= 2048)
            hipLaunchKernelGGL(select_probes_wave_kernel<32>, gw, dim3(256), 0, st, p->logits.as<float>(), p->ncls, mass,
                               np_min, np_max, B, probes, counts);
        else if (p->ncls <= 4096)
            hipLaunchKernelGGL(select_probes_wave_kernel<64>, gw, dim3(256), 0, st, p->logits.as<float>(), p->ncls, mass,
                               np_min, np_max, B, probes, counts);
        else
            hipLaunchKernelGGL(select_probes_kernel, dim3((unsigned)B), dim3(256), 0, st, p->logits.as<float>(),
                               p->ncls, mass, np_min, np_max, probes, counts);
        PN_HIP(hipGetLastError());
    });
}
