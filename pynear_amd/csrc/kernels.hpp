// Host-side launchers of the HIP kernels (one translation unit each) used by the index code.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pn {

// Candidate filter of the multi-index-hashing search (Hamming only): rows hold their sub-strings
// in groups of `group_words` 32-bit words (1 or 2; 0 = no filter), and a row is a candidate when
// one of its first `subs` groups is within `rsub` bits of the query's group. Other rows are never
// returned (their slots come back as index -1).
struct SubFilter {
    int group_words = 0;
    int subs = 0;
    int rsub = 0;
};

// Arguments of a sliced exact scan. The database and the queries are device arrays with a row
// stride of `row_elems` elements (32-bit words for Hamming, floats for the float metrics).
// part: [slices][K][nq_pad] u64 keys (see device_util.hpp), written by the scan.
struct ScanArgs {
    const void *db;
    int64_t n;
    int64_t row_elems;
    int64_t dim;             // real row length (floats) for the float metrics
    const void *queries;
    int64_t nq;
    int64_t rows_per_slice;
    int64_t slices;
    uint64_t *part;
    int64_t nq_pad;
    int qpl = 1;             // queries per lane (Hamming scan)
    SubFilter filt = {};     // Hamming scan: candidate filter (group_words > 0 needs qpl = 1)
};

// Distance kinds for the select stage's output.
enum OutKind { OUT_HAMMING_I32 = 0, OUT_L2_F32 = 1, OUT_F32 = 2 };

// Hamming scan over rows of `words` 32-bit words; K in {1,2,4,8,16,32}; a.qpl in {1,2}
// (nq_pad must be a multiple of 256 * qpl).
void hamming_scan(const ScanArgs &a, int words, int K, hipStream_t st);
// Padded row width (words) served by the Hamming scan for `nbytes`-byte rows; 0 when too wide.
int hamming_padded_words(int64_t nbytes);

// Float scan (reference summation order) for dim <= kFloatScanMaxDim; K in {1,...,32}.
constexpr int kFloatScanMaxDim = 64;
void float_scan(const ScanArgs &a, int metric, int K, hipStream_t st);

// Merge the per-slice lists into kk results per query, ascending (distance, index).
// out_idx int32[nq, kk]; out_dist int32 or float32 [nq, kk] per `kind`.
void merge_lists(const uint64_t *part, int64_t slices, int K, int64_t nq, int64_t nq_pad, int kk, int kind,
                 int32_t *out_idx, void *out_dist, hipStream_t st);

// Wave-per-query merge regardless of the tuning table: per-slice lists of klist keys, per-lane
// capacity cap (1, 2, 4, ..., 64). Exact when cap >= kk (or cap >= klist with one slice per lane).
// Entry j of slice s's list of query q is part[q qs + (s klist + j) es].
void merge_lists_wave(const uint64_t *part, int64_t slices, int klist, int cap, int64_t nq, int64_t qs, int64_t es,
                      int kk, int kind, int32_t *out_idx, void *out_dist, hipStream_t st);

// Exact k-NN through a full distance matrix and a segmented radix sort; any k, any width.
// metric: PN_METRIC_*; rows/queries padded to `row_elems` (words for Hamming, floats otherwise).
struct MatrixSearch {
    void *ws = nullptr;     // workspace, caller-owned
    size_t ws_bytes = 0;
};
size_t matrix_search_workspace(int64_t n, int64_t q_chunk);
void matrix_search(int metric, const void *db, int64_t n, int64_t row_elems, int64_t dim, const void *queries,
                   int64_t nq, int kk, int32_t *out_idx, void *out_dist, void *ws, size_t ws_bytes,
                   int64_t q_chunk, hipStream_t st, const SubFilter &filt = SubFilter());

// Hamming range search over rows of `words` 32-bit words (1, 2, 4, ..., 32), queries padded the
// same way: every (row, distance) with distance <= threshold, per query in ascending row order.
// Pass one keeps up to kRangeChunk matches per (slice, query) and counts all of them; a second
// scan writes the matches straight to their place when a count exceeds that.
constexpr int kRangeChunk = 8;
struct RangePlan {
    int64_t slices, rows_per_slice, nq_pad;
};
// chunk: u32 [slices][kRangeChunk][nq_pad]; cnt, sp: u32 [slices][nq_pad]; tot, mx: u32 [nq]
// (matches of each query; its largest per-slice count). words = 0 takes rows of any width
// (`row_words` words, a slower scan).
void range_count(const void *db, int64_t n, int words, int64_t row_words, const void *queries, int64_t nq,
                 uint32_t threshold, const RangePlan &p, uint32_t *chunk, uint32_t *cnt, uint32_t *sp, uint32_t *tot,
                 uint32_t *mx, hipStream_t st);
// qoff: int64 [nq] start of each query's matches; out_idx / out_dist: int32 [total]
void range_write(const void *db, int64_t n, int words, int64_t row_words, const void *queries, int64_t nq,
                 uint32_t threshold, const RangePlan &p, bool rescan, const uint32_t *chunk, const uint32_t *cnt,
                 const uint32_t *sp, const int64_t *qoff, int32_t *out_idx, int32_t *out_dist, hipStream_t st);

// Row-wise distances out[i] = dist(a[i], b[i]); device pointers.
void pairwise_distance(int metric, const void *a, const void *b, int64_t n, int64_t dim, void *out, hipStream_t st);

} // namespace pn
