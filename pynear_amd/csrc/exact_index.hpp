// Internal interface of the exact index, shared with the IVF indexes (coarse assignment and
// k-means assignment are exact nearest-centroid searches).
#pragma once
#include "common.hpp"
#include "kernels.hpp"
#include "l2_mfma.hpp"

struct pn_exact {
    int32_t metric = PN_METRIC_L2;
    int device = 0;
    int64_t n = 0, dim = 0;      // rows; row length in floats (float metrics) or bytes (Hamming)
    int64_t row_elems = 0;       // padded row length on the device (floats or 32-bit words)
    bool scan_ok = false;        // the sliced scan kernel serves this row width
    uint64_t generation = 0;     // bumped by every set(): plans that bound the rows check it
    pn::DevBuf db, qbuf, part, oidx, odist, ws;
    pn::DevBuf rws, rout;        // range search: counts and chunks; matches
    int cus = 256;
    pn::L2Set l2;                // bf16 split of the rows (L2 metric)
    pn::L2Work l2w;
};


namespace pn {

int64_t padded_row_elems(int32_t metric, int64_t dim, bool *scan_ok);
size_t row_bytes_in(int32_t metric, int64_t dim);
// host rows (in_b bytes each) -> zero-padded device rows (out_b bytes each)
void upload_rows_padded(const void *rows, int64_t n, size_t in_b, size_t out_b, void *dst, hipStream_t st);
// device rows (in_b bytes each) -> zero-padded device rows (out_b bytes each)
void pad_rows_device(const void *src, int64_t n, size_t in_b, size_t out_b, void *dst, hipStream_t st);
void exact_init(pn_exact *h, int32_t metric);
// rows_padded: device rows already padded to padded_row_elems(metric, dim) 32-bit elements
void exact_set_device(pn_exact *h, const void *rows_padded, int64_t n, int64_t dim, hipStream_t st);
// nq padded device queries; out_idx int32[nq, kk]; out_dist int32 (Hamming) or float32 [nq, kk].
// filt (Hamming): only candidate rows are returned; missing results are index -1, distance -1.
void exact_search_enqueue(pn_exact *h, const void *dq, int64_t nq, int kk, int32_t *oi, void *od, hipStream_t st,
                          const SubFilter &filt = SubFilter());
// Sliced-scan shape for a Hamming range search of nq queries over the index's rows.
RangePlan range_plan(const pn_exact *h, int64_t nq);

} // namespace pn
