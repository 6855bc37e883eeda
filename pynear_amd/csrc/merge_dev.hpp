// Device body of the wave-per-query merge of per-slice candidate lists (shared by the single and
// the batched searches).
#pragma once
#include "common.hpp"
#include "device_util.hpp"
#include "kernels.hpp"

namespace pn {

// One wave per query q: lane l merges the per-slice lists (klist keys each) of slices l, l+64, ...
// into a private sorted list of capacity K, then the wave extracts the kk smallest keys by kk
// rounds of a wave-wide minimum (keys are unique since indices are). The result is the exact kk
// smallest of all lists when K >= min(kk, klist * slices per lane).
template <int K>
__device__ __forceinline__ void merge_wave_query(const uint64_t *__restrict__ part, int64_t slices, int klist,
                                                 int64_t nq_pad, int64_t q, int kk, int kind,
                                                 int32_t *__restrict__ out_idx, void *__restrict__ out_dist) {
    const int lane = threadIdx.x & 63;
    uint64_t L[K];
#pragma unroll
    for (int j = 0; j < K; ++j) L[j] = kKey64None;
    for (int64_t s = lane; s < slices; s += 64) {
        const uint64_t *src = part + (s * klist) * nq_pad + q;
        // a slice's list is sorted: keys are read 8 at a time (independent loads), and the list
        // is done at its first empty key or the first key the private list cannot take
        bool done = false;
        for (int j0 = 0; j0 < klist && !done; j0 += 8) {
            uint64_t kb[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) kb[u] = j0 + u < klist ? src[(int64_t)(j0 + u) * nq_pad] : kKey64None;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                uint64_t key = kb[u];
                if (done || key == kKey64None) {
                    done = true;
                    continue;
                }
                if (kind == OUT_L2_F32) {
                    float dd = sqrtf(key_f32((uint32_t)(key >> 32)));
                    key = ((uint64_t)f32_key(dd) << 32) | (key & 0xFFFFFFFFull);
                }
                if (key >= L[K - 1]) {
                    done = true;
                    continue;
                }
                list_insert_u64<K>(L, key);
            }
        }
    }
    for (int r = 0; r < kk; ++r) {
        uint64_t m = L[0];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            uint64_t o = __shfl_xor(m, off, 64);
            m = o < m ? o : m;
        }
        if (L[0] == m && m != kKey64None) { // the unique owner pops its head
#pragma unroll
            for (int j = 0; j < K - 1; ++j) L[j] = L[j + 1];
            L[K - 1] = kKey64None;
        }
        if (lane == 0) {
            out_idx[q * kk + r] = m == kKey64None ? -1 : (int32_t)(uint32_t)(m & 0xFFFFFFFFull);
            uint32_t hi = (uint32_t)(m >> 32);
            if (kind == OUT_HAMMING_I32)
                ((int32_t *)out_dist)[q * kk + r] = (int32_t)hi;
            else
                ((float *)out_dist)[q * kk + r] = key_f32(hi);
        }
    }
}

} // namespace pn
