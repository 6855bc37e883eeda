/*
 * pynear_amd C-ABI -- the drop-in boundary of the MI355X search path.
 *
 * The reference's boundary is the pybind11 module `_pynear` (pynear/src/PythonBindings.cpp:331-481).
 * This header exports the same operations as plain C entry points (plain pointers and sizes, no
 * Python or torch types) so that any binding -- ctypes, pybind11, cgo, JNI -- can sit on top.
 * Each entry point names the reference interface it replaces.
 *
 * Conventions
 *  - Every entry point returns a pn_status; on failure pn_last_error() (thread-local) holds the
 *    message. PN_EINVAL maps to Python ValueError, PN_ETYPE to TypeError, every other non-zero
 *    status to RuntimeError (the exception types pybind11 produces for the reference's throws).
 *  - "host" pointers are ordinary CPU memory, borrowed for the duration of the call; `set` copies
 *    the rows into device memory owned by the handle (the reference copies into the index too,
 *    VPTree.hpp:96-104).
 *  - "_device" entry points take device pointers and a hipStream_t (passed as void*; NULL is the
 *    device's null stream, as everywhere in HIP) and do not synchronise: results are ready when
 *    the stream is.
 *  - Result order is ascending by (distance, index); the Python layer reverses it where the
 *    reference emits farthest-first (VPTree.hpp:621-631).
 *  - Calls on one handle are serialised by the caller (the reference holds the GIL throughout);
 *    handles are independent.
 */
#ifndef PYNEAR_AMD_H
#define PYNEAR_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t pn_status;
enum {
    PN_OK = 0,
    PN_EINVAL = 1,    /* bad argument value or shape  -> ValueError   */
    PN_ERUNTIME = 2,  /* invalid state (empty index)  -> RuntimeError */
    PN_ETYPE = 3,     /* wrong element type           -> TypeError    */
    PN_EDEVICE = 4,   /* HIP runtime failure          -> RuntimeError */
    PN_ENOMEM = 5     /* device allocation failure    -> RuntimeError */
};

typedef enum {
    PN_METRIC_L2 = 0,        /* sqrt(sum (a-b)^2)   DistanceFunctions.hpp:304-345 */
    PN_METRIC_L1 = 1,        /* sum |a-b|           DistanceFunctions.hpp:347-382 */
    PN_METRIC_CHEBYSHEV = 2, /* max |a-b|           DistanceFunctions.hpp:384-420 */
    PN_METRIC_HAMMING = 3    /* popcount(a xor b)   DistanceFunctions.hpp:431-462 */
} pn_metric;

/* Thread-local message of the last failed call on this thread. */
const char *pn_last_error(void);

/* Library version string. */
const char *pn_version(void);

/* Number of visible HIP devices (0 when none). */
pn_status pn_device_count(int32_t *count);

/* Measurement hook (no counterpart in the reference, which has no instrumentation): when enabled,
 * every kernel stage records a pair of HIP events on its launch stream. pn_stage_times() waits for
 * them and returns, per stage (0 scan, 1 select/merge, 2 coarse assignment, 3 rerank, 4 other),
 * the summed milliseconds and launch counts since the previous call, then clears them. */
pn_status pn_stage_timing(int32_t enable);
pn_status pn_stage_times(double *ms /*[5]*/, int64_t *launches /*[5]*/);

/* Launch-shape tuning parameters (no counterpart in the reference): named integers such as
 * "hamming.wg_per_cu" in the library's table (defaults measured on MI355X, overridable through
 * the PN_TUNE environment variable "name=value,..."); pn_batch_tune keeps its own measured
 * schedules per batch shape. Unknown names return PN_EINVAL. */
pn_status pn_tune_set(const char *name, int64_t value);
pn_status pn_tune_get(const char *name, int64_t *value);

/* ---------------------------------------------------------------- distances
 * Row-wise distances out[i] = dist(a[i], b[i]) for n row pairs of `dim` elements (float32 for
 * L2/L1/Chebyshev, bytes for Hamming). Replaces the free functions dist_l2 / dist_l1 /
 * dist_chebyshev / dist_hamming_{64,128,256,512} (PythonBindings.cpp:281-303, :332-338).
 * out is float32[n] for float metrics, int64[n] for Hamming. Host pointers. */
pn_status pn_pairwise_distance(int32_t metric, const void *a, const void *b, int64_t n, int64_t dim, void *out);

/* ---------------------------------------------------------------- exact index
 * Brute-force exact index; replaces the search side of vptree::VPTree<T, dist_t, dist>
 * (VPTree.hpp:92-191) behind VPTreeL2Index / VPTreeL1Index / VPTreeChebyshevIndex /
 * VPTreeBinaryIndex{64,128,256,512,} (PythonBindings.cpp:344-406). */
typedef struct pn_exact pn_exact;

/* VPTreeNumpyAdapter() (PythonBindings.cpp:33). */
pn_status pn_exact_create(int32_t metric, pn_exact **out);
void pn_exact_destroy(pn_exact *h);

/* VPTree::set (VPTree.hpp:92-110) via the adapters' set (PythonBindings.cpp:36-47, :152).
 * rows: host, n x dim (float32 elements, or bytes for Hamming), C-contiguous. n == 0 empties. */
pn_status pn_exact_set(pn_exact *h, const void *rows, int64_t n, int64_t dim);

/* Number of rows and row width of the indexed set (0, 0 when empty). */
pn_status pn_exact_info(const pn_exact *h, int64_t *n, int64_t *dim);

/* VPTree::searchKNN (VPTree.hpp:143-166) / search1NN (:169-191, k = 1).
 * queries: host nq x dim. Writes kk = min(k, n) columns per query, ascending (dist, idx):
 * out_idx int64[nq, kk]; out_dist float32[nq, kk] (float metrics, L2 as sqrt) or int64[nq, kk]
 * (Hamming). An empty index returns PN_ERUNTIME (VPTree.hpp:145-147). */
pn_status pn_exact_search(pn_exact *h, const void *queries, int64_t nq, int64_t dim, int32_t k,
                          int64_t *out_idx, void *out_dist);

/* Device-resident form of pn_exact_search: queries, out_idx (int32[nq, kk]) and out_dist
 * (float32 or int32 [nq, kk]) are device pointers; enqueued on `stream`, no host sync. */
pn_status pn_exact_search_device(pn_exact *h, const void *queries, int64_t nq, int32_t k, int32_t *out_idx,
                                 void *out_dist, void *stream);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): pn_exact_search_device
 * with no allocation at all. Its scratch (per-slice lists, zero-padded queries) is a fixed 32 MiB
 * device array that each GPU holds for the life of the process, so calls that overlap on one GPU
 * race, whatever their handles and streams. Serves k <= 32 for the rows the sliced scan takes:
 * Hamming rows of up to 128 bytes, L2 / L1 / Chebyshev rows of up to 64 floats; other shapes, and
 * query counts whose lists exceed the scratch, return PN_EINVAL. Results are identical to
 * pn_exact_search_device's. */
pn_status pn_exact_search_device_noalloc(pn_exact *h, const void *queries, int64_t nq, int32_t k, int32_t *out_idx,
                                         void *out_dist, void *stream);

/* BKTree::find_batch (BKTree.hpp:117-135) behind BKTreeBinaryIndex{64,128,256,512,}.find_threshold
 * (PythonBindings.cpp:209-230, :408-451) on a Hamming index: every row within `threshold` of each
 * query (dist <= threshold), in compressed rows. queries: host nq x dim bytes. offsets int64[nq + 1]
 * and *total are always written; when *total <= capacity, out_idx / out_dist (int32[*total]) hold
 * the matches of query q at [offsets[q], offsets[q + 1]), in ascending row order (the reference
 * lists the same pairs in its tree's breadth-first order). When *total > capacity nothing else is
 * written: call again with capacity >= *total. An empty index or a negative threshold finds nothing
 * (BKTree.hpp:79-81); a metric other than Hamming returns PN_EINVAL. */
pn_status pn_exact_range(pn_exact *h, const void *queries, int64_t nq, int64_t dim, int64_t threshold,
                         int64_t *offsets, int32_t *out_idx, int32_t *out_dist, int64_t capacity, int64_t *total);

/* ---------------------------------------------------------------- multi-index hashing
 * MIHBinaryIndex (MIH.hpp:50-195) behind the pybind11 class of the same name
 * (PythonBindings.cpp:253-272, :468-481). */
typedef struct pn_mih pn_mih;

/* MIHBinaryIndex(m) (MIH.hpp:57). m >= 1. */
pn_status pn_mih_create(int32_t m, pn_mih **out);
void pn_mih_destroy(pn_mih *h);

/* MIHBinaryIndex::set (MIH.hpp:60-83): host rows n x nbytes, replacing the content; n == 0 clears.
 * nbytes must be a multiple of m, at least m, with nbytes / m <= 8 (PN_EINVAL otherwise, as the
 * reference's invalid_argument); on the GPU also m <= 32 for sub-strings of up to 4 bytes and
 * m <= 16 for 5 to 8 bytes (PN_EINVAL beyond). */
pn_status pn_mih_set(pn_mih *h, const uint8_t *rows, int64_t n, int64_t nbytes);

/* m(), n(), nbytes() (MIH.hpp:154-156). */
pn_status pn_mih_info(const pn_mih *h, int32_t *m, int64_t *n, int64_t *nbytes);

/* MIHBinaryIndex::searchKNN (MIH.hpp:98-152): the kk = min(k, n) nearest candidates of each query,
 * a candidate being a row with one of its m sub-strings within radius / m bits of the query's
 * (the keys the reference enumerates, :102-122), ascending (distance, row). out_idx / out_dist
 * int64[nq, kk] (slots past out_count[q] are -1). radius < 0 and a width other than the indexed
 * one return PN_EINVAL; an empty index returns PN_ERUNTIME. */
pn_status pn_mih_search(pn_mih *h, const uint8_t *queries, int64_t nq, int64_t nbytes, int32_t k, int32_t radius,
                        int64_t *out_idx, int64_t *out_dist, int64_t *out_count);

/* ---------------------------------------------------------------- k-means
 * kmeans_l2(data, k, max_iter, seed) (PythonBindings.cpp:305-329, KMeans.hpp:25-126): host data
 * n x d float32; writes labels int32[n] and centroids float32[k, d]. The result is identical to
 * the reference's (same random draws, same arithmetic). k outside [1, n] returns PN_ERUNTIME. */
pn_status pn_kmeans_l2(const float *data, int64_t n, int64_t d, int64_t k, int64_t max_iter, uint32_t seed,
                       int32_t *labels, float *centroids);

/* ---------------------------------------------------------------- IVF binary index
 * IVFFlatBinaryIndex (BinaryIVF.hpp:34-227) behind the pybind11 class of the same name
 * (PythonBindings.cpp:231-251, :454-466). The build reproduces the reference's k-means++ and
 * Lloyd iterations with the same random draws, so centroids and inverted lists are identical. */
typedef struct pn_ivfbin pn_ivfbin;

/* IVFFlatBinaryIndex(nlist, nprobe, max_iter, seed) (BinaryIVF.hpp:40-46). */
pn_status pn_ivfbin_create(int32_t nlist, int32_t nprobe, int32_t max_iter, uint32_t seed, pn_ivfbin **out);
void pn_ivfbin_destroy(pn_ivfbin *h);

/* IVFFlatBinaryIndex::set (BinaryIVF.hpp:49-60): host rows n x nbytes; n == 0 clears. */
pn_status pn_ivfbin_set(pn_ivfbin *h, const uint8_t *rows, int64_t n, int64_t nbytes);

/* Rows, row width and number of clusters built (min(nlist, n)). */
pn_status pn_ivfbin_info(const pn_ivfbin *h, int64_t *n, int64_t *nbytes, int32_t *n_clusters);

/* nlist() / nprobe() / set_nprobe() (BinaryIVF.hpp:113-115). */
pn_status pn_ivfbin_get_params(const pn_ivfbin *h, int32_t *nlist, int32_t *nprobe);
pn_status pn_ivfbin_set_nprobe(pn_ivfbin *h, int32_t nprobe);

/* Copies the built centroids (n_clusters x nbytes) and/or the final labels (n), either may be
 * NULL. No counterpart in the reference (its build state is private); used to check the build. */
pn_status pn_ivfbin_export(const pn_ivfbin *h, uint8_t *centroids, int32_t *labels);

/* IVFFlatBinaryIndex::searchKNN (BinaryIVF.hpp:65-111): host queries nq x nbytes; per query
 * out_count[q] <= k results, ascending (dist, id), in out_idx/out_dist [nq, k] (rest -1).
 * An empty index yields empty results, as in the reference. k <= 2048. */
pn_status pn_ivfbin_search(pn_ivfbin *h, const uint8_t *queries, int64_t nq, int64_t nbytes, int32_t k,
                           int64_t *out_idx, int64_t *out_dist, int32_t *out_count);

/* Device-resident form: device queries (nq x nbytes), int32 outputs [nq, k] and counts [nq]. */
pn_status pn_ivfbin_search_device(pn_ivfbin *h, const void *queries, int64_t nq, int32_t k, int32_t *out_idx,
                                  int32_t *out_dist, int32_t *out_count, void *stream);

/* ---------------------------------------------------------------- IVF L2 index
 * The search of IVFFlatL2Index (pynear/forest.py:25-115), which is pure numpy in the reference:
 * coarse top-nprobe clusters, then the probed clusters' rows; results ascending by squared L2
 * (forest.py:107-109). Clusters without rows are left out, like the reference's per-cluster
 * arrays (forest.py:51-59). */
typedef struct pn_ivfl2 pn_ivfl2;

pn_status pn_ivfl2_create(pn_ivfl2 **out);
void pn_ivfl2_destroy(pn_ivfl2 *h);

/* IVFFlatL2Index.set (forest.py:36-59) with its k-means (forest.py:136-143 -> kmeans_l2):
 * host rows n x d; runs kmeans_l2(rows, n_clusters, max_iter, seed) on the device, returns its
 * labels (int32[n]) and centroids (float32[n_clusters, d]) and builds the device index. */
pn_status pn_ivfl2_build(pn_ivfl2 *h, const float *rows, int64_t n, int64_t d, int64_t n_clusters, int64_t max_iter,
                         uint32_t seed, int32_t *labels_out, float *centroids_out);

/* Builds the device index from given centroids (nc x d) and row labels (in [0, nc)). */
pn_status pn_ivfl2_set(pn_ivfl2 *h, const float *rows, int64_t n, int64_t d, const float *centroids, int64_t nc,
                       const int32_t *labels);

/* Rows, dimension and number of non-empty clusters. */
pn_status pn_ivfl2_info(const pn_ivfl2 *h, int64_t *n, int64_t *d, int32_t *n_lists);

/* _IVFFlatIndex.searchKNN (forest.py:63-111): host queries nq x d; nprobe clusters per query
 * (capped at the non-empty cluster count); any k. For k <= 64, precision 0: bf16 candidate
 * distances, 1: three-product bf16 (~fp32) candidate distances, 2: fp8 e4m3 candidate distances
 * on the residuals to the list centroid (one power-of-two scale; d <= 1152); all are followed by
 * an fp32 squared-L2 rerank of max(4 k, 40) (at most 64) candidates. For k > 64 every row of the
 * probed lists is scored in fp32 with the rerank's arithmetic (precision does not apply) and the
 * k smallest are kept. out_idx int64[nq, k], out_dist float32[nq, k] (squared L2), out_count. */
pn_status pn_ivfl2_search(pn_ivfl2 *h, const float *queries, int64_t nq, int64_t d, int32_t nprobe, int32_t k,
                          int32_t precision, int64_t *out_idx, float *out_dist, int32_t *out_count);

/* Device-resident form: device queries (nq x d, contiguous), int32/float32 outputs [nq, k];
 * k <= 64 (larger k: pn_ivfl2_search). */
pn_status pn_ivfl2_search_device(pn_ivfl2 *h, const float *queries, int64_t nq, int32_t nprobe, int32_t k,
                                 int32_t precision, int32_t *out_idx, float *out_dist, int32_t *out_count,
                                 void *stream);

/* The list scan and rerank of pn_ivfl2_search_device for probes chosen by the caller (no coarse
 * stage; used with a learned probe predictor): probes int32[nq, np_max] are k-means cluster ids
 * (ids of clusters without rows, and slots j >= counts[q], are skipped). No reference counterpart:
 * the reference always probes the n_probe nearest centroids (forest.py:76-80). */
pn_status pn_ivfl2_search_probes_device(pn_ivfl2 *h, const float *queries, int64_t nq, const int32_t *probes,
                                        const int32_t *counts, int32_t np_max, int32_t k, int32_t precision,
                                        int32_t *out_idx, float *out_dist, int32_t *out_count, void *stream);

/* Work of the last fp8 list scan on this handle (precision 2): the matrix-core instructions it
 * issued (v_mfma_scale_f32_32x32x64_f8f6f4, 131072 flops each) and the bytes of fp8 rows its work
 * items read (rows that both query groups of an item scan counted once), after group pruning.
 * Read after the search's stream has completed. No reference counterpart: measurement only (the
 * roofline of the scan, whose work depends on the data). */
pn_status pn_ivfl2_scan_work(pn_ivfl2 *h, int64_t *mfma_ops, int64_t *row_bytes);

/* ---------------------------------------------------------------- learned probe predictor
 * An MLP d -> hidden (ReLU) -> nlist in bf16 (fp32 accumulation and masters) that ranks a query's
 * IVF clusters, trained on the device on exact neighbours; its softmax picks an adaptive number
 * of probes for pn_ivfl2_search_probes_device. No reference counterpart (the reference probes
 * the n_probe nearest centroids, forest.py:76-80). Parameters are one fp32 vector
 * [W1 (hidden x d) | b1 (hidden) | W2 (nlist x hidden) | b2 (nlist)]. d, hidden and nlist must be
 * multiples of 16; training batches too. */
typedef struct pn_probe pn_probe;

/* He-initialised weights (std::mt19937(seed)), zero biases, Adam state zero. */
pn_status pn_probe_create(int32_t d, int32_t hidden, int32_t nlist, uint32_t seed, pn_probe **out);
void pn_probe_destroy(pn_probe *p);
pn_status pn_probe_get_params(const pn_probe *p, float *params);
pn_status pn_probe_set_params(pn_probe *p, const float *params);

/* Forward + backward on device inputs x (B x d fp32) and labels (B x L cluster ids: the clusters of
 * each query's exact top-L neighbours; the target puts 1/L on each). Writes the batch-mean
 * cross-entropy to *loss_out and the gradient (parameter layout) to grads_out (device; either may
 * be NULL). */
pn_status pn_probe_grads_device(pn_probe *p, const float *x, const int32_t *labels, int32_t L, int64_t B,
                                float *loss_out, float *grads_out, void *stream);

/* One Adam step (beta 0.9 / 0.999, eps 1e-8, learning rate lr) on the batch's gradient. */
pn_status pn_probe_train_step_device(pn_probe *p, const float *x, const int32_t *labels, int32_t L, int64_t B,
                                     float lr, float *loss_out, void *stream);

/* logits_out (device, B x nlist fp32) for device queries x (B x d). */
pn_status pn_probe_forward_device(pn_probe *p, const float *x, int64_t B, float *logits_out, void *stream);

/* Adaptive probes: per query the clusters of highest predicted probability, the fewest whose mass
 * reaches `mass`, between np_min and np_max; probes int32[B x np_max], counts int32[B] (device). */
pn_status pn_probe_select_device(pn_probe *p, const float *x, int64_t B, float mass, int32_t np_min, int32_t np_max,
                                 int32_t *probes, int32_t *counts, void *stream);

/* ---------------------------------------------------------------- batched exact search
 * Many exact L2 searches over independent indices of uneven sizes served together. No single
 * call of the reference corresponds: each sub-search has the semantics of pn_exact_search_device
 * on its index (VPTree::searchKNN, VPTree.hpp:143-166) and returns bit-identical results. */
typedef struct pn_batch pn_batch;

/* Plans n sub-searches: indices[i] (exact L2 indices on one device), device queries[i]
 * (nq[i] x dim_i float32, contiguous), 1 <= k[i] <= 32, device outputs out_idx[i]
 * (int32[nq[i], min(k[i], n_i)]) and out_dist[i] (float32, sqrt L2, same shape). The plan binds
 * these pointers and captures the search (one split, one persistent scan over all tiles of all
 * indices, one merge, rerank and fallback launch) in a HIP graph; its schedule comes from the
 * tuning table (batch.slice_rows, batch.wg_per_cu, batch.order). The plan also binds each index's
 * device rows as they are now: after pn_exact_set on any of its indices, pn_batch_run and
 * pn_batch_tune return PN_EINVAL (plan a new batch); the indices must outlive the plan. */
pn_status pn_batch_create(int32_t n, pn_exact *const *indices, const void *const *queries, const int64_t *nq,
                          const int32_t *k, int32_t *const *out_idx, float *const *out_dist, pn_batch **out);
void pn_batch_destroy(pn_batch *b);

/* Replays the plan's graph on `stream` (NULL = the null stream); no host synchronisation. */
pn_status pn_batch_run(pn_batch *b, void *stream);

/* Autotuner: times a set of schedules (slice length, persistent grid size, tile order; `reps`
 * replays each) and re-plans with the fastest. With table_path, a schedule stored there for
 * this batch shape is used without measuring, and a measured one is appended. */
pn_status pn_batch_tune(pn_batch *b, const char *table_path, int32_t reps);

/* Number of scan tiles and the schedule in use: cfg = {slice_rows, wg_per_cu, order}. */
pn_status pn_batch_info(const pn_batch *b, int64_t *tiles, int32_t *cfg);

#ifdef __cplusplus
}
#endif
#endif /* PYNEAR_AMD_H */
