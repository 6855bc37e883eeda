// Device body of the wave-per-query merge of per-slice candidate lists (shared by the single and
// the batched searches).
#pragma once
#include "common.hpp"
#include "device_util.hpp"
#include "kernels.hpp"

namespace pn {

// One wave per query q: lane l merges the per-slice lists (klist keys each) of slices l, l+64, ...
// Entry j of slice s's list of query q is part[q qs + (s klist + j) es] (query-major: qs =
// slices klist, es = 1; list-major: qs = 1, es = the padded query count).
// into a private sorted list of capacity K, then the wave orders what the lanes hold and writes
// the kk smallest keys (keys are unique since indices are). The result is the exact kk smallest
// of all lists when K >= min(kk, klist * slices per lane).
//
// A sampled bound (merge_sample_bound) first limits what the lanes insert: the kk <= 64 smallest
// entries are all at or below it.
//
// Ordering: when the lanes hold at most 256 keys together, they are packed into the wave's 256
// entries of LDS (wbuf) and sorted by one bitonic network across the wave; otherwise kk rounds of
// a wave-wide minimum pop them one by one.
//
// kind OUT_L2_F32 orders by sqrt of the stored value: a slice stays sorted in the high word, but
// entries whose square roots tie may follow in any index order, so a slice is left only at a key
// whose high word exceeds the limit's.
// OUT_L2_F32 lists hold squared values; the merge orders by their square roots (high word), the
// row in the low word kept.
__device__ __forceinline__ uint64_t merge_transform(uint64_t key, bool tr) {
    if (tr && key != kKey64None) {
        const float dd = sqrtf(key_f32((uint32_t)(key >> 32)));
        key = ((uint64_t)f32_key(dd) << 32) | (key & 0xFFFFFFFFull);
    }
    return key;
}

// Sampled bound of the wave merges: at least min(kk, all entries) entries lie at or below it.
// Lane l samples up to four entries at grid positions l + 64 u (slice g % slices, entry
// g / slices: the first entries of every slice). For kk <= 16 the bound is the wave maximum of
// the lanes' minima, with at least 64 distinct entries at or below it (one per lane); for larger
// kk it is the kk-th smallest sample (the 256 sorted across the wave). Missing samples are empty
// keys (the largest): a wave with too few samples leaves the bound open.
__device__ __forceinline__ uint64_t merge_sample_bound(const uint64_t *__restrict__ part, int64_t slices, int klist,
                                                       int64_t qs, int64_t es, int64_t q, int kk, bool tr, int lane) {
    const int G = (int)min<int64_t>(slices * klist, 256), ns = (int)min<int64_t>(slices, 256);
    uint64_t smp[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int g = lane + 64 * u;
        const int gs = g < G ? g : 0;
        const int64_t sl = gs % ns, p = gs / ns;
        const uint64_t v = part[q * qs + (sl * klist + p) * es];
        smp[u] = g < G ? v : kKey64None;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) smp[u] = merge_transform(smp[u], tr);
    if (kk > 16) {
        wave_sort_u64<4, 2>(smp, lane);
        return __shfl(smp[0], kk - 1, 64); // (kk <= 64: rank kk - 1 is v[0] of lane kk - 1)
    }
    uint64_t bnd = kKey64None;
#pragma unroll
    for (int u = 0; u < 4; ++u) bnd = smp[u] < bnd ? smp[u] : bnd;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t o = __shfl_xor(bnd, off, 64);
        bnd = o > bnd ? o : bnd;
    }
    return bnd;
}

// KEEP: nothing is written; lane r < kk receives the r-th smallest key in `mine` (others: empty).
template <int K, bool KEEP = false>
__device__ __forceinline__ void merge_wave_query(const uint64_t *__restrict__ part, int64_t slices, int klist,
                                                 int64_t qs, int64_t es, int64_t q, int kk, int kind,
                                                 int32_t *__restrict__ out_idx, void *__restrict__ out_dist,
                                                 uint64_t *wbuf, uint64_t &mine) {
    const int lane = threadIdx.x & 63;
    const bool tr = kind == OUT_L2_F32;
    auto transform = [&](uint64_t key) { return merge_transform(key, tr); };
    auto write_out = [&](int r, uint64_t m) {
        out_idx[q * kk + r] = m == kKey64None ? -1 : (int32_t)(uint32_t)(m & 0xFFFFFFFFull);
        const uint32_t hi = (uint32_t)(m >> 32);
        if (kind == OUT_HAMMING_I32)
            ((int32_t *)out_dist)[q * kk + r] = (int32_t)hi;
        else
            ((float *)out_dist)[q * kk + r] = key_f32(hi);
    };
    const uint64_t bnd = merge_sample_bound(part, slices, klist, qs, es, q, kk, tr, lane);
    if (KEEP) mine = kKey64None;
    uint64_t L[K];
#pragma unroll
    for (int j = 0; j < K; ++j) L[j] = kKey64None;
    for (int64_t s = lane; s < slices; s += 64) {
        const uint64_t *src = part + q * qs + (s * klist) * es;
        // a slice's list is sorted: keys are read 8 at a time (independent loads, clamped to the
        // list), and the list is done at its first empty key or the first key past the limit
        // (the bound or the private list's last key)
        bool done = false;
        for (int j0 = 0; j0 < klist && !done; j0 += 8) {
            uint64_t kb[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) kb[u] = src[(int64_t)min(j0 + u, klist - 1) * es];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (done || j0 + u >= klist || kb[u] == kKey64None) {
                    done = true;
                    continue;
                }
                const uint64_t key = transform(kb[u]);
                const uint64_t lim = bnd < L[K - 1] ? bnd : L[K - 1];
                if (key > bnd || key >= L[K - 1]) {
                    if (!tr || (key >> 32) > (lim >> 32)) done = true;
                    continue;
                }
                list_insert_u64<K>(L, key);
            }
        }
    }
    int c = 0; // keys this lane holds
#pragma unroll
    for (int j = 0; j < K; ++j) c += L[j] != kKey64None ? 1 : 0;
    int off = c; // inclusive prefix over lanes
#pragma unroll
    for (int dd = 1; dd < 64; dd <<= 1) {
        const int t = __shfl_up(off, dd, 64);
        if (lane >= dd) off += t;
    }
    const int total = __shfl(off, 63, 64);
    off -= c;
    if (total <= 256) {
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (j < c) wbuf[off + j] = L[j];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint64_t v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = lane + 64 * u < total ? wbuf[lane + 64 * u] : kKey64None;
        wave_sort_u64<4, 2>(v, lane);
        if (KEEP) mine = lane < kk ? v[0] : kKey64None;
        else if (lane < kk) write_out(lane, v[0]); // (kk <= 64: the kk smallest are v[0] of lanes 0 .. kk-1)
        return;
    }
    for (int r = 0; r < kk; ++r) {
        uint64_t m = L[0];
#pragma unroll
        for (int sh = 32; sh > 0; sh >>= 1) {
            uint64_t o = __shfl_xor(m, sh, 64);
            m = o < m ? o : m;
        }
        if (L[0] == m && m != kKey64None) { // the unique owner pops its head
#pragma unroll
            for (int j = 0; j < K - 1; ++j) L[j] = L[j + 1];
            L[K - 1] = kKey64None;
        }
        if (KEEP) {
            if (lane == r) mine = m;
        } else if (lane == 0) {
            write_out(r, m);
        }
    }
}

// Wave-per-query merge without private lists (any kk <= 64; entry layout as merge_wave_query).
// The kk smallest keys lie at or below a bound b that has at least kk keys at or below it. Each
// lane counts the keys at or below b in its slices (a slice is sorted: reading stops at its first
// key past b; under OUT_L2_F32 past b's high word, as ties of square roots may come in any index
// order) and keeps its four smallest. With at most 256 such keys in the wave they are packed into
// the wave's LDS (wbuf; offsets by a prefix sum of the counts; a lane with more than four re-reads
// its lists) and sorted by one bitonic network. Otherwise b is lowered to the kk-th smallest of
// the lanes' four smallest (kk keys lie at or below it) and the count repeats, twice at most; a
// wave that still holds more than 256 keys pops the kk smallest in rounds of a wave minimum over
// the lanes' four smallest, a lane refilling them from its lists when they run out. rounds_only
// takes the rounds at once (a test switch for that path). KEEP as merge_wave_query.
template <bool KEEP = false>
__device__ __forceinline__ void merge_wave_lean(const uint64_t *__restrict__ part, int64_t slices, int klist,
                                                int64_t qs, int64_t es, int64_t q, int kk, int kind,
                                                int32_t *__restrict__ out_idx, void *__restrict__ out_dist,
                                                uint64_t *wbuf, bool rounds_only, uint64_t &mine) {
    const int lane = threadIdx.x & 63;
    const bool tr = kind == OUT_L2_F32;
    auto write_out = [&](int r, uint64_t m) {
        out_idx[q * kk + r] = m == kKey64None ? -1 : (int32_t)(uint32_t)(m & 0xFFFFFFFFull);
        const uint32_t hi = (uint32_t)(m >> 32);
        if (kind == OUT_HAMMING_I32)
            ((int32_t *)out_dist)[q * kk + r] = (int32_t)hi;
        else
            ((float *)out_dist)[q * kk + r] = key_f32(hi);
    };
    // f(key) for every key of this lane's slices in (lo, b] (lo = kKey64None: no lower limit)
    auto visit = [&](uint64_t b, uint64_t lo, auto &&f) {
        for (int64_t s = lane; s < slices; s += 64) {
            const uint64_t *src = part + q * qs + (s * klist) * es;
            bool done = false;
            for (int j0 = 0; j0 < klist && !done; j0 += 8) {
                uint64_t kb[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) kb[u] = src[(int64_t)min(j0 + u, klist - 1) * es];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (done || j0 + u >= klist || kb[u] == kKey64None) {
                        done = true;
                        continue;
                    }
                    const uint64_t key = merge_transform(kb[u], tr);
                    if (key > b) {
                        if (!tr || (key >> 32) > (b >> 32)) done = true;
                        continue;
                    }
                    if (lo == kKey64None || key > lo) f(key);
                }
            }
        }
    };
    uint64_t bnd = merge_sample_bound(part, slices, klist, qs, es, q, kk, tr, lane);
    uint64_t L4[4];
    int c = 0, off = 0, total = 0;
    for (int it = 0;; ++it) {
        c = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) L4[j] = kKey64None;
        visit(bnd, kKey64None, [&](uint64_t key) {
            ++c;
            if (key < L4[3]) list_insert_u64<4>(L4, key);
        });
        off = c; // inclusive prefix over lanes
#pragma unroll
        for (int dd = 1; dd < 64; dd <<= 1) {
            const int t = __shfl_up(off, dd, 64);
            if (lane >= dd) off += t;
        }
        total = __shfl(off, 63, 64);
        off -= c;
        if (rounds_only) {
            total = 257;
            break;
        }
        if (total <= 256 || it == 2) break;
        uint64_t v[4] = {L4[0], L4[1], L4[2], L4[3]};
        wave_sort_u64<4, 2>(v, lane);
        const uint64_t nb = __shfl(v[0], kk - 1, 64);
        if (nb >= bnd) break; // no progress (the keys sit in a few lanes): rounds
        bnd = nb;
    }
    if (total <= 256) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (c <= 4 && j < c) wbuf[off + j] = L4[j];
        if (__any(c > 4)) {
            int i = 0;
            if (c > 4) visit(bnd, kKey64None, [&](uint64_t key) { wbuf[off + i++] = key; });
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint64_t v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = lane + 64 * u < total ? wbuf[lane + 64 * u] : kKey64None;
        wave_sort_u64<4, 2>(v, lane);
        if (KEEP) mine = lane < kk ? v[0] : kKey64None;
        else if (lane < kk) write_out(lane, v[0]); // (kk <= 64: the kk smallest are v[0] of lanes 0 .. kk-1)
        return;
    }
    if (KEEP) mine = kKey64None;
    // rounds: L4 holds the lane's four smallest keys at or below bnd; `more` marks keys beyond them
    bool more = c > 4;
    for (int r = 0; r < kk; ++r) {
        uint64_t m = L4[0];
#pragma unroll
        for (int sh = 32; sh > 0; sh >>= 1) {
            const uint64_t o = __shfl_xor(m, sh, 64);
            m = o < m ? o : m;
        }
        if (L4[0] == m && m != kKey64None) { // the unique owner pops its head
#pragma unroll
            for (int j = 0; j < 3; ++j) L4[j] = L4[j + 1];
            L4[3] = kKey64None;
            if (L4[0] == kKey64None && more) { // refill: the four smallest above the popped key
                int cnt = 0;
                visit(bnd, m, [&](uint64_t key) {
                    ++cnt;
                    if (key < L4[3]) list_insert_u64<4>(L4, key);
                });
                more = cnt > 4;
            }
        }
        if (KEEP) {
            if (lane == r) mine = m;
        } else if (lane == 0) {
            write_out(r, m);
        }
    }
}

} // namespace pn
