// pn_exact: brute-force exact index behind the VPTree*Index classes (C-ABI in pynear_amd.h).
//
// Replaces vptree::VPTree<T, dist_t, dist>::set/searchKNN/search1NN (VPTree.hpp:92-191) for the
// float adapters (PythonBindings.cpp:33-145) and the binary adapters (:147-200). The tree walk
// is replaced by a full scan; the result contract is kept: min(n, k) neighbours per query, exact
// distances in the reference's arithmetic, an error on an empty index (VPTree.hpp:145-147).
//
// Device layout: rows padded to `row_elems` elements (32-bit words for Hamming, floats for the
// float metrics; zero padding is neutral in every metric), contiguous, row-major.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "device_util.hpp"
#include "kernels.hpp"
#include "exact_index.hpp"

namespace pn {

int64_t padded_row_elems(int32_t metric, int64_t dim, bool *scan_ok) {
    if (metric == PN_METRIC_HAMMING) {
        int w = hamming_padded_words(dim);
        *scan_ok = w > 0;
        return w > 0 ? w : ceil_div(dim, 4);
    }
    *scan_ok = dim <= kFloatScanMaxDim;
    if (!*scan_ok) return dim;
    for (int p : {8, 16, 32, 64})
        if (dim <= p) return p;
    return dim;
}

static size_t elem_bytes(int32_t metric) { return 4; }

size_t row_bytes_in(int32_t metric, int64_t dim) {
    return metric == PN_METRIC_HAMMING ? (size_t)dim : (size_t)dim * sizeof(float);
}

void upload_rows_padded(const void *rows, int64_t n, size_t in_b, size_t out_b, void *dst, hipStream_t st) {
    if (in_b == out_b) {
        PN_HIP(hipMemcpyAsync(dst, rows, in_b * n, hipMemcpyHostToDevice, st));
        return;
    }
    std::vector<uint8_t> tmp(out_b * n, 0);
    for (int64_t i = 0; i < n; ++i) std::memcpy(tmp.data() + i * out_b, (const uint8_t *)rows + i * in_b, in_b);
    PN_HIP(hipMemcpyAsync(dst, tmp.data(), tmp.size(), hipMemcpyHostToDevice, st));
    PN_HIP(hipStreamSynchronize(st)); // tmp is released on return
}

// Copies n host rows of `dim` elements into a zero-padded device array of `row_elems` elements.
static void upload_padded(const pn_exact *h, const void *rows, int64_t n, int64_t dim, void *dst, hipStream_t st) {
    upload_rows_padded(rows, n, row_bytes_in(h->metric, dim), (size_t)h->row_elems * elem_bytes(h->metric), dst, st);
}

void pad_rows_device(const void *src, int64_t n, size_t in_b, size_t out_b, void *dst, hipStream_t st) {
    if (in_b == out_b) {
        PN_HIP(hipMemcpyAsync(dst, src, in_b * n, hipMemcpyDeviceToDevice, st));
        return;
    }
    PN_HIP(hipMemsetAsync(dst, 0, out_b * n, st));
    PN_HIP(hipMemcpy2DAsync(dst, out_b, src, in_b, in_b, n, hipMemcpyDeviceToDevice, st));
}

void exact_init(pn_exact *h, int32_t metric) {
    h->metric = metric;
    h->device = current_device();
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) == hipSuccess && cus > 0)
        h->cus = cus;
}

void exact_set_device(pn_exact *h, const void *rows_padded, int64_t n, int64_t dim, hipStream_t st) {
    bool scan_ok = false;
    int64_t re = padded_row_elems(h->metric, dim, &scan_ok);
    ++h->generation;
    h->dim = dim;
    h->row_elems = re;
    h->scan_ok = scan_ok;
    void *dst = h->db.ensure((size_t)std::max<int64_t>(n, 1) * re * 4);
    if (n) PN_HIP(hipMemcpyAsync(dst, rows_padded, (size_t)n * re * 4, hipMemcpyDeviceToDevice, st));
    if (h->metric == PN_METRIC_L2) l2_prepare(h->l2, (const float *)dst, n, dim, re, st);
    h->n = n;
}

static int list_size_for(int kk) {
    int K = 1;
    while (K < kk) K <<= 1;
    return K;
}

static int out_kind_of(int32_t metric) {
    return metric == PN_METRIC_HAMMING ? OUT_HAMMING_I32 : (metric == PN_METRIC_L2 ? OUT_L2_F32 : OUT_F32);
}

// Shape of a sliced scan: slices S of rps rows each, queries padded to nq_pad (whole 256 * qpl
// groups). With part_cap > 0 the per-slice lists (S * K * nq_pad u64 keys) must fit part_cap
// bytes: S is lowered as far as the slice length limit allows; false when even that does not fit.
struct SlicePlan {
    int64_t S, rps, nq_pad;
    int qpl;
};
static bool plan_slices(const pn_exact *h, int64_t nq, int K, size_t part_cap, SlicePlan &p, bool one_qpl = false) {
    p.qpl = (!one_qpl && h->metric == PN_METRIC_HAMMING && tune_get("hamming.qpl") == 2 && nq > 256) ? 2 : 1;
    const int64_t gx = ceil_div(nq, 256 * p.qpl);
    p.nq_pad = gx * 256 * p.qpl;
    // Whole multiples of the CU count: every CU holds the same number of resident workgroups.
    const int64_t wpc = tune_get(h->metric == PN_METRIC_HAMMING ? "hamming.wg_per_cu" : "float.wg_per_cu");
    const int64_t target = std::max<int64_t>(1, wpc) * h->cus;
    int64_t S = std::max<int64_t>(1, target / gx);
    S = std::min<int64_t>(S, std::max<int64_t>(1, h->n / 64));
    const int64_t s_min = ceil_div(h->n, kMaxRowsPerSlice);
    if (part_cap) {
        const int64_t per_slice = (int64_t)K * p.nq_pad * (int64_t)sizeof(uint64_t);
        S = std::min<int64_t>(S, (int64_t)(part_cap / (size_t)per_slice));
        if (S < s_min) return false;
    }
    S = std::max<int64_t>(S, s_min);
    p.rps = ceil_div(h->n, S);
    p.S = ceil_div(h->n, p.rps);
    return true;
}

// The sliced scan (Hamming and the float metrics up to kFloatScanMaxDim dims, K <= 32) and the
// merge of its per-slice lists; part holds the plan's lists.
static void sliced_search(const pn_exact *h, const void *dq, int64_t nq, int kk, int K, const SlicePlan &p,
                          uint64_t *part, int32_t *oi, void *od, hipStream_t st, const SubFilter &filt = SubFilter()) {
    ScanArgs a{h->db.p, h->n, h->row_elems, h->dim, dq, nq, p.rps, p.S, part, p.nq_pad, p.qpl, filt};
    {
        StageScope sc(st, STAGE_SCAN);
        if (h->metric == PN_METRIC_HAMMING)
            hamming_scan(a, (int)h->row_elems, K, st);
        else
            float_scan(a, h->metric, K, st);
    }
    StageScope sc(st, STAGE_SELECT);
    merge_lists(part, p.S, K, nq, p.nq_pad, kk, out_kind_of(h->metric), oi, od, st);
}

RangePlan range_plan(const pn_exact *h, int64_t nq) {
    SlicePlan p;
    plan_slices(h, nq, 1, 0, p, true);
    return {p.S, p.rps, p.nq_pad};
}

// Enqueues the search of nq padded device queries; out_idx int32[nq, kk], out_dist int32/f32.
void exact_search_enqueue(pn_exact *h, const void *dq, int64_t nq, int kk, int32_t *oi, void *od, hipStream_t st,
                          const SubFilter &filt) {
    const bool filtered = filt.group_words > 0;
    PN_REQUIRE(!filtered || h->metric == PN_METRIC_HAMMING, PN_EINVAL, "the candidate filter is for Hamming rows");
    int K = list_size_for(kk);
    if (filtered) K = kk <= 4 ? 4 : (kk <= 16 ? 16 : (kk <= 32 ? 32 : 64)); // the filtered scan's list sizes
    if (h->metric == PN_METRIC_L2 && kk <= 32 && tune_get("l2.mfma") != 0) {
        l2_mfma_search(h->l2, (const float *)h->db.p, h->row_elems, (const float *)dq, h->row_elems, nq, kk, oi,
                       (float *)od, h->l2w, h->cus, st);
        return;
    }
    if (h->scan_ok && K <= 32 && (!filtered || h->row_elems % filt.group_words == 0)) {
        SlicePlan p;
        plan_slices(h, nq, K, 0, p, filtered);
        uint64_t *part = (uint64_t *)h->part.ensure((size_t)p.S * K * p.nq_pad * sizeof(uint64_t));
        sliced_search(h, dq, nq, kk, K, p, part, oi, od, st, filt);
        return;
    }
    // Distance-matrix path: bound the matrix to ~1 GiB of keys per chunk.
    int64_t q_chunk = std::max<int64_t>(1, std::min<int64_t>({nq, (int64_t)(256 << 20) / std::max<int64_t>(h->n, 1),
                                                                (int64_t)(INT32_MAX - 1) / std::max<int64_t>(h->n, 1),
                                                                60000}));
    size_t wsb = matrix_search_workspace(h->n, q_chunk);
    void *ws = h->ws.ensure(wsb);
    StageScope sc(st, STAGE_OTHER);
    matrix_search(h->metric, h->db.p, h->n, h->row_elems, h->dim, dq, nq, kk, oi, od, ws, wsb, q_chunk, st, filt);
}

} // namespace pn

using namespace pn;

extern "C" pn_status pn_exact_create(int32_t metric, pn_exact **out) {
    return guard([&] {
        PN_REQUIRE(out, PN_EINVAL, "out is null");
        PN_REQUIRE(metric >= PN_METRIC_L2 && metric <= PN_METRIC_HAMMING, PN_EINVAL, "unknown metric");
        auto *h = new pn_exact();
        exact_init(h, metric);
        *out = h;
    });
}

extern "C" void pn_exact_destroy(pn_exact *h) {
    if (!h) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(h->device);
    delete h;
    if (prev >= 0) (void)hipSetDevice(prev);
}

extern "C" pn_status pn_exact_set(pn_exact *h, const void *rows, int64_t n, int64_t dim) {
    return guard([&] {
        PN_REQUIRE(h, PN_EINVAL, "index is null");
        PN_REQUIRE(n >= 0 && dim >= 0, PN_EINVAL, "negative shape");
        DeviceScope ds(h->device);
        ++h->generation; // (whatever follows: the old rows are gone)
        if (n == 0) { // VPTree::set on an empty array leaves the index empty (VPTree.hpp:93-94)
            h->n = 0;
            h->dim = 0;
            h->db.release();
            return;
        }
        PN_REQUIRE(rows, PN_EINVAL, "rows is null");
        PN_REQUIRE(dim > 0, PN_EINVAL, "rows must have at least one element");
        PN_REQUIRE(n < (int64_t)INT32_MAX, PN_EINVAL, "too many rows (indices are 32-bit on the device)");
        bool scan_ok = false;
        int64_t re = padded_row_elems(h->metric, dim, &scan_ok);
        h->n = 0;
        h->dim = dim;
        h->row_elems = re;
        h->scan_ok = scan_ok;
        hipStream_t st = lib_stream();
        void *dst = h->db.ensure((size_t)n * re * elem_bytes(h->metric));
        upload_padded(h, rows, n, dim, dst, st);
        if (h->metric == PN_METRIC_L2) l2_prepare(h->l2, (const float *)dst, n, dim, re, st);
        PN_HIP(hipStreamSynchronize(st));
        h->n = n;
    });
}

extern "C" pn_status pn_exact_info(const pn_exact *h, int64_t *n, int64_t *dim) {
    return guard([&] {
        PN_REQUIRE(h && n && dim, PN_EINVAL, "null argument");
        *n = h->n;
        *dim = h->dim;
    });
}

extern "C" pn_status pn_exact_search(pn_exact *h, const void *queries, int64_t nq, int64_t dim, int32_t k,
                                     int64_t *out_idx, void *out_dist) {
    return guard([&] {
        PN_REQUIRE(h, PN_EINVAL, "index is null");
        PN_REQUIRE(h->n > 0, PN_ERUNTIME, "index must be first initialized with .set() function and non empty dataset");
        PN_REQUIRE(nq >= 0, PN_EINVAL, "negative query count");
        PN_REQUIRE(k >= 0, PN_EINVAL, "k must be non-negative");
        PN_REQUIRE(dim == h->dim, PN_EINVAL,
                   "invalid data dimension: index built data and query data dimensions must agree");
        int kk = (int)std::min<int64_t>(k, h->n);
        if (nq == 0 || kk == 0) return;
        PN_REQUIRE(queries && out_idx && out_dist, PN_EINVAL, "null pointer");
        DeviceScope ds(h->device);
        hipStream_t st = lib_stream();
        size_t qb = (size_t)nq * h->row_elems * 4;
        void *dq = h->qbuf.ensure(qb);
        upload_padded(h, queries, nq, dim, dq, st);
        int32_t *oi = (int32_t *)h->oidx.ensure((size_t)nq * kk * 4);
        void *od = h->odist.ensure((size_t)nq * kk * 4);
        exact_search_enqueue(h, dq, nq, kk, oi, od, st);
        std::vector<int32_t> hi((size_t)nq * kk), hd((size_t)nq * kk);
        PN_HIP(hipMemcpyAsync(hi.data(), oi, hi.size() * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipMemcpyAsync(hd.data(), od, hd.size() * 4, hipMemcpyDeviceToHost, st));
        PN_HIP(hipStreamSynchronize(st));
        for (size_t i = 0; i < hi.size(); ++i) out_idx[i] = hi[i];
        if (h->metric == PN_METRIC_HAMMING) {
            int64_t *o = (int64_t *)out_dist;
            for (size_t i = 0; i < hd.size(); ++i) o[i] = hd[i];
        } else {
            std::memcpy(out_dist, hd.data(), hd.size() * 4);
        }
    });
}

extern "C" pn_status pn_exact_search_device(pn_exact *h, const void *queries, int64_t nq, int32_t k,
                                            int32_t *out_idx, void *out_dist, void *stream) {
    return guard([&] {
        PN_REQUIRE(h, PN_EINVAL, "index is null");
        PN_REQUIRE(h->n > 0, PN_ERUNTIME, "index must be first initialized with .set() function and non empty dataset");
        PN_REQUIRE(nq >= 0 && k >= 0, PN_EINVAL, "negative size");
        int kk = (int)std::min<int64_t>(k, h->n);
        if (nq == 0 || kk == 0) return;
        PN_REQUIRE(queries && out_idx && out_dist, PN_EINVAL, "null pointer");
        DeviceScope ds(h->device);
        hipStream_t st = (hipStream_t)stream;
        const void *dq = queries;
        size_t in_b = row_bytes_in(h->metric, h->dim);
        size_t pad_b = (size_t)h->row_elems * 4;
        if (in_b != pad_b) { // pad into the workspace (zero tail per row)
            void *tmp = h->qbuf.ensure((size_t)nq * pad_b);
            pad_rows_device(queries, nq, in_b, pad_b, tmp, st);
            dq = tmp;
        }
        exact_search_enqueue(h, dq, nq, kk, out_idx, out_dist, st);
    });
}

// ---------------------------------------------------------------- search without allocation
// pn_exact_search_device_noalloc keeps its per-slice lists and padded queries in `scratch`, a
// fixed device array (each GPU holds its own copy for the life of the process); its device
// address on each GPU is taken once with hipGetSymbolAddress and used from then on.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

static void *noalloc_scratch(int dev) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void *ptrs[kMaxDevices] = {};
    PN_REQUIRE(dev >= 0 && dev < kMaxDevices, PN_EDEVICE, "device id out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!ptrs[dev]) { // (the caller has made dev current)
        void *p = nullptr;
        PN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PN_REQUIRE(p, PN_EDEVICE, "no device address for the search scratch");
        ptrs[dev] = p;
    }
    return ptrs[dev];
}

// Rows of in_b bytes -> rows of out_b bytes (a multiple of 4, >= in_b) with a zero tail; one
// thread per 32-bit output word. (A kernel, not a memset, so the call can sit in a captured graph.)
__global__ void pad_rows_kernel(const uint8_t *__restrict__ src, int64_t n, int64_t in_b, int64_t out_w,
                                uint32_t *__restrict__ dst) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * out_w) return;
    const int64_t r = t / out_w, b0 = 4 * (t - r * out_w);
    const uint8_t *s = src + r * in_b;
    uint32_t v = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (b0 + j < in_b) v |= (uint32_t)s[b0 + j] << (8 * j);
    dst[t] = v;
}

extern "C" pn_status pn_exact_search_device_noalloc(pn_exact *h, const void *queries, int64_t nq, int32_t k,
                                                    int32_t *out_idx, void *out_dist, void *stream) {
    return guard([&] {
        PN_REQUIRE(h, PN_EINVAL, "index is null");
        PN_REQUIRE(h->n > 0, PN_ERUNTIME, "index must be first initialized with .set() function and non empty dataset");
        PN_REQUIRE(nq >= 0 && k >= 0, PN_EINVAL, "negative size");
        const int kk = (int)std::min<int64_t>(k, h->n);
        if (nq == 0 || kk == 0) return;
        PN_REQUIRE(queries && out_idx && out_dist, PN_EINVAL, "null pointer");
        const int K = list_size_for(kk);
        PN_REQUIRE(h->scan_ok && K <= 32, PN_EINVAL,
                   "search without allocation serves k <= 32 and rows the sliced scan takes (Hamming rows of the "
                   "scan's widths, float rows of at most 64 dims)");
        DeviceScope ds(h->device);
        hipStream_t st = (hipStream_t)stream;
        uint8_t *base = (uint8_t *)noalloc_scratch(h->device);
        const size_t in_b = row_bytes_in(h->metric, h->dim);
        const size_t pad_b = (size_t)h->row_elems * 4;
        const size_t qbytes = in_b != pad_b ? (nq * pad_b + 255) / 256 * 256 : 0; // padded queries first
        PN_REQUIRE(qbytes < kNoAllocScratchBytes, PN_EINVAL, "queries exceed the 32 MiB search scratch");
        SlicePlan p;
        PN_REQUIRE(plan_slices(h, nq, K, kNoAllocScratchBytes - qbytes, p), PN_EINVAL,
                   "query count exceeds the 32 MiB search scratch of the allocation-free search");
        const void *dq = queries;
        if (qbytes) {
            const int64_t words = (int64_t)(pad_b / 4);
            hipLaunchKernelGGL(pad_rows_kernel, dim3((unsigned)ceil_div(nq * words, (int64_t)256)), dim3(256), 0, st,
                               (const uint8_t *)queries, nq, (int64_t)in_b, words, (uint32_t *)base);
            PN_HIP(hipGetLastError());
            dq = base;
        }
        sliced_search(h, dq, nq, kk, K, p, (uint64_t *)(base + qbytes), out_idx, out_dist, st);
        PN_HIP(hipGetLastError());
    });
}

extern "C" pn_status pn_pairwise_distance(int32_t metric, const void *a, const void *b, int64_t n, int64_t dim,
                                          void *out) {
    return guard([&] {
        PN_REQUIRE(metric >= PN_METRIC_L2 && metric <= PN_METRIC_HAMMING, PN_EINVAL, "unknown metric");
        PN_REQUIRE(n >= 0 && dim >= 0, PN_EINVAL, "negative shape");
        if (n == 0) return;
        PN_REQUIRE(a && b && out, PN_EINVAL, "null pointer");
        hipStream_t st = lib_stream();
        size_t rb = row_bytes_in(metric, dim) * n;
        size_t ob = n * (metric == PN_METRIC_HAMMING ? 8 : 4);
        DevBuf da, db, dout;
        da.ensure(rb);
        db.ensure(rb);
        dout.ensure(ob);
        PN_HIP(hipMemcpyAsync(da.p, a, rb, hipMemcpyHostToDevice, st));
        PN_HIP(hipMemcpyAsync(db.p, b, rb, hipMemcpyHostToDevice, st));
        pairwise_distance(metric, da.p, db.p, n, dim, dout.p, st);
        PN_HIP(hipMemcpyAsync(out, dout.p, ob, hipMemcpyDeviceToHost, st));
        PN_HIP(hipStreamSynchronize(st));
    });
}
