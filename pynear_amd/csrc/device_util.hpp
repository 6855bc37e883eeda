// Device-side helpers shared by the scan and select kernels.
//
// Ordering used everywhere: a candidate is the pair (distance, index) compared lexicographically,
// which makes the selection a total order (lower index wins a distance tie). Inside a scan wave
// the pair is packed into one unsigned key so that a single unsigned compare implements it:
//   Hamming  u32 lane key  = d << 21 | row-in-slice         (d <= 1024 for rows up to 1024 bits)
//   any      u64 list key  = monotone-bits(distance) << 32 | global index
// Non-negative float distances map to monotone unsigned keys by their IEEE bit pattern.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace pn {

constexpr int kHamIdxBits = 21;                     // row-in-slice bits of the Hamming lane key
constexpr uint32_t kHamIdxMask = (1u << kHamIdxBits) - 1u;
constexpr int64_t kMaxRowsPerSlice = int64_t(1) << kHamIdxBits;
constexpr uint32_t kKey32None = 0xFFFFFFFFu;
constexpr uint64_t kKey64None = ~uint64_t(0);
// Largest grid of the per-query fallback launches (workgroups of 256 take query chunks b,
// b + grid, ...): a grid of one workgroup per query would pass the 2^32 work-item limit at 16.7 M
// queries.
constexpr int64_t kFallbackMaxGrid = int64_t(1) << 20;
// Queries per fallback workgroup: its first kFallbackChunk threads read their queries' flags at
// once and the workgroup then runs the flagged ones in turn, so a launch over many unflagged
// queries (the usual case) is a few workgroups per CU, not one per query.
constexpr int kFallbackChunk = 16;

// Workgroups are dealt to the eight XCDs round robin (workgroup b runs on XCD b % 8). For a grid
// of a multiple of 8 workgroups this numbering gives each XCD one contiguous eighth of the block
// indices, so blocks that read neighbouring columns of an [entry][query] array (16 queries of 8
// bytes per 128-byte line) share one L2 instead of fetching each line into four.
__device__ __forceinline__ int64_t xcd_contiguous_block() {
    const int64_t per = gridDim.x >> 3;
    return (int64_t)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

// Grid size for xcd_contiguous_block: blocks rounded up to a multiple of 8.
inline unsigned xcd_grid(int64_t blocks) { return (unsigned)((blocks + 7) / 8 * 8); }

// A pointer loaded from device memory (a field of a work descriptor) is generic to the compiler,
// whose flat loads and stores count against both the vector-memory and the LDS wait counters, so
// an LDS read then waits for every global load in flight. Every pointer a descriptor holds points
// to global memory: rebuilt through an integer as a global-address-space pointer, the accesses
// through it compile to global loads and stores.
template <class T> __device__ __forceinline__ T *as_global(T *p) {
    return (T *)(__attribute__((address_space(1))) T *)(uintptr_t)p;
}

// Insert `key` into the ascending list L[0..K-1] (L[K-1] is dropped). With L sorted,
// max(L[j-1], min(key, L[j])) == med3(L[j-1], key, L[j]), one v_med3_u32 per slot.
template <int K> __device__ __forceinline__ void list_insert_u32(uint32_t (&L)[K], uint32_t key) {
#pragma unroll
    for (int j = K - 1; j > 0; --j) L[j] = max(L[j - 1], min(key, L[j]));
    L[0] = min(L[0], key);
}

template <int K> __device__ __forceinline__ void list_insert_u64(uint64_t (&L)[K], uint64_t key) {
#pragma unroll
    for (int j = K - 1; j > 0; --j) {
        uint64_t lo = key < L[j] ? key : L[j];
        L[j] = L[j - 1] > lo ? L[j - 1] : lo;
    }
    L[0] = key < L[0] ? key : L[0];
}

// Sorts 64 E keys held E per lane (element lane + 64 s in v[s]) ascending across the wave
// (bitonic network; partners within a lane for strides >= 64, else one lane shuffle away). The
// input is already made of sorted runs of K0 / 2 that alternate ascending / descending, so the
// network starts at blocks of K0.
template <int E, int K0>
__device__ __forceinline__ void wave_sort_u64(uint64_t (&v)[E], int lane) {
    constexpr int N = 64 * E;
#pragma unroll
    for (int k = K0; k <= N; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
#pragma unroll
                for (int s = 0; s < E; ++s) {
                    const int t = s ^ (j / 64);
                    if (t > s) {
                        const bool asc = ((lane + 64 * s) & k) == 0;
                        const uint64_t a = v[s], b = v[t];
                        const bool sw = asc ? a > b : a < b;
                        v[s] = sw ? b : a;
                        v[t] = sw ? a : b;
                    }
                }
            } else {
#pragma unroll
                for (int s = 0; s < E; ++s) {
                    const uint64_t o = __shfl_xor(v[s], j, 64);
                    const bool asc = ((lane + 64 * s) & k) == 0, lower = (lane & j) == 0;
                    const uint64_t mn = o < v[s] ? o : v[s], mx = o < v[s] ? v[s] : o;
                    v[s] = lower == asc ? mn : mx;
                }
            }
        }
}

// Sum over the wave by data-parallel-primitive moves (each step one VALU add reading a permuted
// operand): pairs and quads of lanes, mirrored half rows and rows (sums of 16), then rows 1 and 3
// add the last lane of the row before and rows 2 and 3 that of row 1, so lane 63 ends with the
// total, which every lane takes.
__device__ __forceinline__ float wave_sum_dpp(float v) {
    auto dpp = [](float x, auto ctrl, auto rows) __attribute__((always_inline)) {
        return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x),
                                                                     decltype(ctrl)::value, decltype(rows)::value,
                                                                     0xF, false));
    };
    using full = std::integral_constant<int, 0xF>;
    v += dpp(v, std::integral_constant<int, 0xB1>{}, full{});  // quad_perm [1, 0, 3, 2]
    v += dpp(v, std::integral_constant<int, 0x4E>{}, full{});  // quad_perm [2, 3, 0, 1]
    v += dpp(v, std::integral_constant<int, 0x141>{}, full{}); // row_half_mirror
    v += dpp(v, std::integral_constant<int, 0x140>{}, full{}); // row_mirror
    v += dpp(v, std::integral_constant<int, 0x142>{}, std::integral_constant<int, 0xA>{}); // row_bcast:15
    v += dpp(v, std::integral_constant<int, 0x143>{}, std::integral_constant<int, 0xC>{}); // row_bcast:31
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}

// The same bitonic sort for 32-bit keys (one shuffle per cross-lane step).
template <int E>
__device__ __forceinline__ void wave_sort_u32(uint32_t (&v)[E], int lane) {
    constexpr int N = 64 * E;
#pragma unroll
    for (int k = 2; k <= N; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
#pragma unroll
                for (int s = 0; s < E; ++s) {
                    const int t = s ^ (j / 64);
                    if (t > s) {
                        const bool asc = ((lane + 64 * s) & k) == 0;
                        const uint32_t a = v[s], b = v[t];
                        v[s] = asc ? min(a, b) : max(a, b);
                        v[t] = asc ? max(a, b) : min(a, b);
                    }
                }
            } else {
#pragma unroll
                for (int s = 0; s < E; ++s) {
                    const uint32_t o = (uint32_t)__shfl_xor((int)v[s], j, 64);
                    const bool asc = ((lane + 64 * s) & k) == 0, lower = (lane & j) == 0;
                    v[s] = lower == asc ? min(o, v[s]) : max(o, v[s]);
                }
            }
        }
}

// popcount(x) + acc as one v_bcnt_u32_b32. Written as an (ALU-only) asm statement because the
// compiler otherwise reassociates chains of __popc(x) + acc into bcnt(x, 0) pairs joined by
// v_add3_u32, 25% more VALU work per word.
__device__ __forceinline__ uint32_t bcnt_acc(uint32_t x, uint32_t acc) {
    uint32_t r;
    asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(acc));
    return r;
}

// Ascending bitonic sorting network over 16 keys held in registers (80 compare-exchanges).
template <typename T> __device__ __forceinline__ void sort16(T (&k)[16]) {
#pragma unroll
    for (int size = 2; size <= 16; size <<= 1)
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int j = i ^ stride;
                if (j > i) {
                    const T lo = k[i] < k[j] ? k[i] : k[j], hi = k[i] < k[j] ? k[j] : k[i];
                    const bool asc = (i & size) == 0;
                    k[i] = asc ? lo : hi;
                    k[j] = asc ? hi : lo;
                }
            }
}

__device__ __forceinline__ uint32_t f32_key(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float key_f32(uint32_t u) { return __uint_as_float(u); }

} // namespace pn
