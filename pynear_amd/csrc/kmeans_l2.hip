// kmeans_l2 on the GPU, reproducing the reference's result bit for bit (KMeans.hpp:25-126).
//
//   * k-means++: per new centroid a kernel folds (dist_l2)^2 to the previous centroid into min_sq
//     with the reference's arithmetic (8 FMA chains, one 8-lane group per row, sqrt, then v*v).
//     The host sums min_sq in index order in double and draws unif01 exactly as the reference
//     (std::mt19937 + libstdc++ distributions), then scans the same cumulative sum.
//   * Lloyd assignment: exact nearest centroid through the exact L2 search (k = 1, lower centroid
//     index winning ties like the reference's strict `<`).
//   * update: per cluster and dimension, the member rows are summed in ascending point order
//     (a stable sort of the labels gives that order) starting from 0.f, then scaled by the
//     correctly rounded 1.f / count -- the reference's serial loop, one thread per (cluster, dim).
//     Empty clusters take a pick_point row, in cluster order.
#include <hipcub/hipcub.hpp>

#include <cstring>
#include <limits>
#include <random>
#include <vector>

#include "common.hpp"
#include "exact_index.hpp"
#include "kernels.hpp"
#include "kmeans_l2.hpp"
#include "ref_arith.hpp"

namespace pn {


__global__ __launch_bounds__(256) void kpp_l2_kernel(const float *__restrict__ X, int64_t n, int64_t d, int64_t stride,
                                                     const float *__restrict__ c, float *__restrict__ min_sq) {
    const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 3;
    const int g = threadIdx.x & 7;
    if (row >= n) return; // the 8 lanes of a row leave together
    const float core = l2_core_group8(X + row * stride, c, d, g);
    if (g == 0) {
        float v = sqrtf(core);
        v = v * v;
        if (v < min_sq[row]) min_sq[row] = v;
    }
}

__global__ void fill_f32_kernel(float *p, int64_t n, float v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void iota_kernel(int32_t *p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}

__global__ void km_relabel_kernel(const int32_t *__restrict__ nl, int32_t *__restrict__ ol, int64_t n,
                                  unsigned long long *__restrict__ changed, int32_t *__restrict__ counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool ch = false;
    if (i < n) {
        const int32_t v = nl[i];
        ch = v != ol[i];
        ol[i] = v;
        atomicAdd(&counts[v], 1);
    }
    const unsigned long long b = __ballot(ch);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(changed, (unsigned long long)__popcll(b));
}

// C[c, j] = (sum over members m of c, ascending, of X[m, j]) * (1 / count); empty clusters untouched.
__global__ __launch_bounds__(256) void centroid_sum_kernel(const float *__restrict__ X, int64_t d, int64_t stride,
                                                           const int32_t *__restrict__ members,
                                                           const int32_t *__restrict__ off, float *__restrict__ C) {
    const int c = blockIdx.x;
    const int64_t j = (int64_t)blockIdx.y * blockDim.x + threadIdx.x;
    const int b0 = off[c], b1 = off[c + 1];
    if (j >= d || b1 == b0) return;
    float acc = 0.f;
    for (int m = b0; m < b1; ++m) acc += X[(int64_t)members[m] * stride + j];
    const float inv = 1.f / (float)(b1 - b0);
    C[(int64_t)c * d + j] = acc * inv;
}

void kmeans_l2_device(const float *X, int64_t n, int64_t d, int64_t stride, int64_t k, int64_t max_iter, uint32_t seed,
                      int32_t *labels_out, float *centroids_out, hipStream_t st) {
    PN_REQUIRE(k >= 1 && k <= n, PN_ERUNTIME, "kmeans_l2: k must be between 1 and N");
    PN_REQUIRE(n < (int64_t)INT32_MAX, PN_EINVAL, "kmeans_l2: too many rows");
    std::mt19937 rng(seed);
    std::uniform_int_distribution<size_t> pick(0, (size_t)n - 1);
    std::uniform_real_distribution<float> unif01(0.f, 1.f);
    const size_t rowb = (size_t)d * 4;
    float *C = centroids_out; // k x d, contiguous
    auto copy_row = [&](int64_t dst_c, int64_t src_row) {
        PN_HIP(hipMemcpy2DAsync(C + dst_c * d, rowb, X + src_row * stride, rowb, rowb, 1, hipMemcpyDeviceToDevice, st));
    };

    // k-means++ (KMeans.hpp:37-70)
    copy_row(0, (int64_t)pick(rng));
    DevBuf minsq;
    float *dmin = (float *)minsq.ensure((size_t)n * 4);
    hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, dmin, n,
                       std::numeric_limits<float>::max());
    std::vector<float> hmin(n);
    for (int64_t ci = 1; ci < k; ++ci) {
        hipLaunchKernelGGL(kpp_l2_kernel, dim3((unsigned)ceil_div(n * 8, 256)), dim3(256), 0, st, X, n, d, stride,
                           C + (ci - 1) * d, dmin);
        PN_HIP(hipMemcpyAsync(hmin.data(), dmin, (size_t)n * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipStreamSynchronize(st));
        double total = 0.0;
        for (float v : hmin) total += (double)v;
        const double r = (double)unif01(rng) * total;
        double cum = 0.0;
        int64_t chosen = n - 1;
        for (int64_t i = 0; i < n; ++i) {
            cum += (double)hmin[i];
            if (cum >= r) {
                chosen = i;
                break;
            }
        }
        copy_row(ci, chosen);
    }

    // Lloyd (KMeans.hpp:72-123)
    DevBuf nl, ol, counts, chg, idx_in, keys_out, members, offs, tmp, ddist;
    int32_t *dnl = (int32_t *)nl.ensure((size_t)n * 4);
    int32_t *dol = labels_out;
    int32_t *dcnt = (int32_t *)counts.ensure((size_t)k * 4);
    unsigned long long *dchg = (unsigned long long *)chg.ensure(8);
    int32_t *dio = (int32_t *)idx_in.ensure((size_t)n * 4);
    int32_t *dko = (int32_t *)keys_out.ensure((size_t)n * 4);
    int32_t *dmem = (int32_t *)members.ensure((size_t)n * 4);
    int32_t *doff = (int32_t *)offs.ensure((size_t)(k + 1) * 4);
    void *dd = ddist.ensure((size_t)n * 4);
    int end_bit = 1;
    while ((int64_t(1) << end_bit) < k) ++end_bit;
    size_t tmp_bytes = 0;
    PN_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, (const int32_t *)nullptr, (int32_t *)nullptr,
                                              (const int32_t *)nullptr, (int32_t *)nullptr, (int)n, 0, end_bit, st));
    void *dtmp = tmp.ensure(tmp_bytes);
    PN_HIP(hipMemsetAsync(dol, 0, (size_t)n * 4, st));
    hipLaunchKernelGGL(iota_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, dio, n);
    std::vector<int32_t> hcnt(k), hoff(k + 1);
    pn_exact cent;
    exact_init(&cent, PN_METRIC_L2);
    bool scan_ok = false;
    const int64_t re = padded_row_elems(PN_METRIC_L2, d, &scan_ok);
    DevBuf cpad, xpad;
    const float *Xq = X;
    if (re != stride) { // the exact search reads queries with the padded row stride (zero tail)
        float *xp = (float *)xpad.ensure((size_t)n * re * 4);
        PN_HIP(hipMemsetAsync(xp, 0, (size_t)n * re * 4, st));
        PN_HIP(hipMemcpy2DAsync(xp, (size_t)re * 4, X, (size_t)stride * 4, rowb, n, hipMemcpyDeviceToDevice, st));
        Xq = xp;
    }
    float *cp = (float *)cpad.ensure((size_t)k * re * 4);
    for (int64_t it = 0; it < max_iter; ++it) {
        pad_rows_device(C, k, rowb, (size_t)re * 4, cp, st);
        exact_set_device(&cent, cp, k, d, st);
        exact_search_enqueue(&cent, Xq, n, 1, dnl, dd, st);
        PN_HIP(hipMemsetAsync(dchg, 0, 8, st));
        PN_HIP(hipMemsetAsync(dcnt, 0, (size_t)k * 4, st));
        hipLaunchKernelGGL(km_relabel_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, dnl, dol, n, dchg,
                           dcnt);
        unsigned long long changed = 0;
        PN_HIP(hipMemcpyAsync(&changed, dchg, 8, hipMemcpyDeviceToHost, st));
        PN_HIP(hipMemcpyAsync(hcnt.data(), dcnt, (size_t)k * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipStreamSynchronize(st));
        if (changed == 0) break;
        hoff[0] = 0;
        for (int64_t c = 0; c < k; ++c) hoff[c + 1] = hoff[c] + hcnt[c];
        PN_HIP(hipMemcpyAsync(doff, hoff.data(), (size_t)(k + 1) * 4, hipMemcpyHostToDevice, st));
        size_t tb = tmp_bytes;
        PN_HIP(hipcub::DeviceRadixSort::SortPairs(dtmp, tb, dol, dko, dio, dmem, (int)n, 0, end_bit, st));
        hipLaunchKernelGGL(centroid_sum_kernel, dim3((unsigned)k, (unsigned)ceil_div(d, 256)), dim3(256), 0, st, X, d,
                           stride, dmem, doff, C);
        for (int64_t c = 0; c < k; ++c)
            if (hcnt[c] == 0) copy_row(c, (int64_t)pick(rng));
        PN_HIP(hipStreamSynchronize(st)); // hoff is rewritten next iteration
    }
    PN_HIP(hipGetLastError());
}

} // namespace pn

using namespace pn;

extern "C" pn_status pn_kmeans_l2(const float *data, int64_t n, int64_t d, int64_t k, int64_t max_iter, uint32_t seed,
                                  int32_t *labels, float *centroids) {
    return guard([&] {
        PN_REQUIRE(n >= 0 && d >= 0 && k >= 0 && max_iter >= 0, PN_EINVAL, "negative argument");
        PN_REQUIRE(k >= 1 && k <= n, PN_ERUNTIME, "kmeans_l2: k must be between 1 and N");
        PN_REQUIRE(d >= 1, PN_ERUNTIME, "kmeans_l2: data must be a 2D float32 array (N, D)");
        PN_REQUIRE(data && labels && centroids, PN_EINVAL, "null pointer");
        hipStream_t st = lib_stream();
        DevBuf X, L, C;
        float *dX = (float *)X.ensure((size_t)n * d * 4);
        int32_t *dL = (int32_t *)L.ensure((size_t)n * 4);
        float *dC = (float *)C.ensure((size_t)k * d * 4);
        PN_HIP(hipMemcpyAsync(dX, data, (size_t)n * d * 4, hipMemcpyHostToDevice, st));
        kmeans_l2_device(dX, n, d, d, k, max_iter, seed, dL, dC, st);
        PN_HIP(hipMemcpyAsync(labels, dL, (size_t)n * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipMemcpyAsync(centroids, dC, (size_t)k * d * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipStreamSynchronize(st));
    });
}
