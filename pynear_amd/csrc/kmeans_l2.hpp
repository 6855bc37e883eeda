// Device k-means used by the kmeans_l2 entry point and the IVF-L2 build.
#pragma once
#include "common.hpp"

namespace pn {

// X: n device rows of d floats with row stride `stride`; writes n labels and k x d centroids
// (contiguous) on the device. Reproduces kmeans_l2 (KMeans.hpp:25-126) exactly.
void kmeans_l2_device(const float *X, int64_t n, int64_t d, int64_t stride, int64_t k, int64_t max_iter, uint32_t seed,
                      int32_t *labels_out, float *centroids_out, hipStream_t st);

} // namespace pn
