/*
 * CPU oracle for the pynear_amd search path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library restates the reference's CPU arithmetic so that tests, the smoke check and
 * the bench's cpu_baseline leg can check the HIP path against it. It is never linked into,
 * loaded by, or called from the product (pynear_amd/), which fails loudly without its HIP
 * library instead of falling back to anything here.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * reference checkout root).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* metric ids, shared with include/pynear_amd.h */
enum { OR_L2 = 0, OR_L1 = 1, OR_LINF = 2 };

/* pynear/include/DistanceFunctions.hpp:304-345 (dist_l2_f_avx2), summation order kept. */
float or_dist_l2(const float *a, const float *b, int64_t d);
/* DistanceFunctions.hpp:347-382 (dist_l1_f_avx2) */
float or_dist_l1(const float *a, const float *b, int64_t d);
/* DistanceFunctions.hpp:384-420 (dist_chebyshev_f_avx2) */
float or_dist_linf(const float *a, const float *b, int64_t d);
/* DistanceFunctions.hpp:50-169, :431-462: popcount(a xor b) over nbytes bytes (true count). */
int64_t or_dist_hamming(const uint8_t *a, const uint8_t *b, int64_t nbytes);

/* Brute-force exact k-NN over the metric above; results ascending by (dist, idx);
 * kk = min(k, n) columns are written per query. */
void or_knn_float(int metric, const float *X, int64_t n, int64_t d, const float *Q, int64_t nq,
                  int64_t k, int64_t *out_idx, float *out_dist);
void or_knn_hamming(const uint8_t *X, int64_t n, int64_t nbytes, const uint8_t *Q, int64_t nq,
                    int64_t k, int64_t *out_idx, int64_t *out_dist);

/* pynear/include/BinaryIVF.hpp:127-227 (IVFFlatBinaryIndex::_build). Writes kc = min(nlist, n)
 * centroids and n labels; returns kc. */
int32_t or_ivfbin_build(const uint8_t *X, int64_t n, int64_t nbytes, int32_t nlist, int32_t max_iter,
                        uint32_t seed, uint8_t *centroids, int32_t *labels);
/* BinaryIVF.hpp:65-111 (IVFFlatBinaryIndex::searchKNN) given centroids and inverted lists in CSR
 * form (list_off[nc+1], list_ids ascending within a list). out_* are [nq, k], padded with -1;
 * out_count[q] = number of valid entries. */
void or_ivfbin_search(const uint8_t *cent, int32_t nc, const int64_t *list_off, const int32_t *list_ids,
                      const uint8_t *X, int64_t nbytes, const uint8_t *Q, int64_t nq, int32_t nprobe,
                      int64_t k, int64_t *out_idx, int64_t *out_dist, int32_t *out_count);

/* pynear/include/KMeans.hpp:25-126 (kmeans_l2). Writes n labels and k*d centroids. */
void or_kmeans_l2(const float *X, int64_t n, int64_t d, int64_t k, int64_t max_iter, uint32_t seed,
                  int32_t *labels, float *centroids);

#ifdef __cplusplus
}
#endif
