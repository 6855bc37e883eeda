// Device bodies of the exact L2 search (bf16 MFMA candidate scan, exact rerank, fallback), shared
// by the single-index kernels of l2_mfma.hip and the batched kernels of batch.hip.
#pragma once
#include "common.hpp"
#include "device_util.hpp"
#include "l2_mfma.hpp"
#include "ref_arith.hpp"

namespace pn {

// bf16 hi/lo split and fp64-accumulated squared norm of one row by one wave; rows [n, n_out)
// are written as zero rows (padding that later kernels read unguarded).
__device__ __forceinline__ void split_row(const float *__restrict__ X, int64_t n, int64_t d, int64_t stride, int dp,
                                          __bf16 *__restrict__ H, __bf16 *__restrict__ Lo, float *__restrict__ nrm,
                                          int64_t row) {
    const int lane = threadIdx.x & 63;
    double acc = 0.0;
    if ((d & 3) == 0 && (stride & 3) == 0 && ((uintptr_t)X & 15) == 0) {
        // four consecutive columns per lane: 16-byte row loads, 8-byte stores of each plane (dp is a
        // multiple of 16; the norm's sum runs in another order, which fp64 makes immaterial)
        for (int e = 4 * lane; e < dp; e += 256) {
            const float4 x4 = e < d && row < n ? *(const float4 *)(X + row * stride + e) : make_float4(0.f, 0.f, 0.f, 0.f);
            const float xs[4] = {x4.x, x4.y, x4.z, x4.w};
            union { __bf16 b[4]; uint2 u; } hh, ll;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                hh.b[u] = (__bf16)xs[u];
                ll.b[u] = (__bf16)(xs[u] - (float)hh.b[u]);
                acc += (double)xs[u] * (double)xs[u];
            }
            *(uint2 *)(H + row * dp + e) = hh.u;
            *(uint2 *)(Lo + row * dp + e) = ll.u;
        }
    } else {
        for (int e = lane; e < dp; e += 64) {
            const float x = e < d && row < n ? X[row * stride + e] : 0.f;
            const __bf16 h = (__bf16)x;
            const __bf16 l = (__bf16)(x - (float)h);
            H[row * dp + e] = h;
            Lo[row * dp + e] = l;
            acc += (double)x * (double)x;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) nrm[row] = (float)acc;
}


// Keys of one 32 x 32 tile for this lane's query (rows t + 8g + 4 half + {0..3}, g < 4) into
// the lane's sorted list (Lk, Li) of KL. A key is the approximation's float bits with the low bits
// (mask lowm) replaced by the row: keys stay unique, the minimum names its row, and (key & ~lowm)
// is a lower bound of the approximation. With KO = false the low 5 bits hold the row's offset in
// its tile and Li the rows; with KO ("keys only", slices of at most lowm + 1 rows) they hold the
// row's offset in its slice, the row is r0 + (key & lowm), and the lists carry no rows at all, so
// an insertion costs one three-way median per entry instead of five operations. Passing keys
// (below thr) are inserted in ascending order; thr may come from the other half-wave's list of the
// same query or from the query's other slices: rows at or above it cannot reach the KL best.
// Merges 16 ascending keys b (of the tile at row t) into the ascending list (Lk, Li) of KL = 16 or
// 32, keeping the KL smallest: the list followed by b reversed is bitonic, so one half-cleaner
// step (entry i takes the smaller of Lk[i] and the key facing it) leaves the KL smallest as a
// bitonic sequence, which a bitonic merge sorts. KL - 1 + KL log2(KL) / 2 compare-exchanges.
template <int KL, bool KO = false>
__device__ __forceinline__ void merge_sorted16(const uint32_t (&b)[16], uint32_t t, uint32_t (&Lk)[KL],
                                               uint32_t (&Li)[KL]) {
    static_assert(KL == 16 || KL == 32, "lists of 16 or 32 keys");
#pragma unroll
    for (int i = KL - 16; i < KL; ++i) {
        const uint32_t bk = b[KL - 1 - i];
        if (KO) {
            Lk[i] = min(bk, Lk[i]);
        } else {
            const bool take = bk < Lk[i];
            Li[i] = take ? t + (bk & 31u) : Li[i];
            Lk[i] = take ? bk : Lk[i];
        }
    }
#pragma unroll
    for (int s = KL / 2; s > 0; s >>= 1)
#pragma unroll
        for (int i = 0; i < KL; ++i)
            if ((i & s) == 0) {
                const uint32_t ka = Lk[i], kb = Lk[i + s];
                if (KO) {
                    Lk[i] = min(ka, kb);
                    Lk[i + s] = max(ka, kb);
                } else {
                    const bool sw = kb < ka;
                    const uint32_t ia = Li[i], ib = Li[i + s];
                    Lk[i] = sw ? kb : ka;
                    Lk[i + s] = sw ? ka : kb;
                    Li[i] = sw ? ib : ia;
                    Li[i + s] = sw ? ia : ib;
                }
            }
}

// NET: with many keys to insert, take the network path below (the LDS-staged wide scan; in the
// other scans its registers would spill)
// rb: the slice's first row (KO: rows are numbered from it in the keys); lowm: the row bits.
template <int KL, bool MASK, bool FIRST, bool NET = false, bool KO = false>
__device__ __forceinline__ void tile_keys_insert(const f32x16 &acc, const float4 (&xv)[4], float qnv, int64_t t,
                                                 int64_t r1, int half, uint32_t (&Lk)[KL], uint32_t (&Li)[KL],
                                                 uint32_t &thr, int64_t rb = 0, uint32_t lowm = 31u) {
    float v[16];
    float vmin = INFINITY;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const float xr = (r & 3) == 0 ? xv[r >> 2].x : (r & 3) == 1 ? xv[r >> 2].y : (r & 3) == 2 ? xv[r >> 2].z : xv[r >> 2].w;
        v[r] = fmaxf(0.f, (qnv + xr) - 2.f * acc[r]);
        vmin = fminf(vmin, v[r]);
    }
    // every key of the tile is >= bits(vmin) with the offset bits cleared (masked rows only lower
    // vmin): a tile no lane can take from is done
    if (!__any((f32_key(vmin) & ~lowm) < thr)) return;
    uint32_t key[16];
    uint32_t m1 = kKey32None, m2 = kKey32None; // the lane's two smallest keys (FAST2)
    const uint32_t tb = KO ? (uint32_t)(t - rb) : 0u; // the tile's first row in the key's row bits
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int off = 8 * (r >> 2) + 4 * half + (r & 3);
        key[r] = !MASK || t + off < r1 ? ((f32_key(v[r]) & ~lowm) | (tb + (uint32_t)off)) : kKey32None;
        m2 = min(m2, max(m1, key[r]));
        m1 = min(m1, key[r]);
    }
    // FAST2: for short lists the two smallest keys are found in the pass above and go in without a
    // rescan of the tile's keys; longer lists (whose insertion dominates and whose registers are
    // scarce) take every key by rounds
    constexpr bool FAST2 = KL <= 8;
    if (FIRST) { // empty list: the tile's KL smallest keys, sorted, by a network instead of KL rounds
        sort16(key);
#pragma unroll
        for (int j = 0; j < KL; ++j) {
            Lk[j] = j < 16 ? key[j] : kKey32None;
            if (!KO) Li[j] = j < 16 ? (uint32_t)t + (key[j] & 31u) : 0u;
        }
        thr = min(thr, Lk[KL - 1]);
        return;
    }
    // many keys below thr (a list still filling, as over short slices with k = 32): the tile's keys
    // under thr are sorted and merged by networks in one go -- the same KL smallest as the rounds
    // below take, whose cost grows with the number inserted
    if constexpr (NET && (KL == 16 || KL == 32)) {
        int below = 0;
#pragma unroll
        for (int r = 0; r < 16; ++r) below += key[r] < thr ? 1 : 0;
        if (__any(below >= 4)) {
#pragma unroll
            for (int r = 0; r < 16; ++r) key[r] = key[r] < thr ? key[r] : kKey32None;
            sort16(key);
            merge_sorted16<KL, KO>(key, (uint32_t)t, Lk, Li);
            thr = min(thr, Lk[KL - 1]);
            return;
        }
    }
    // keys below thr go in ascending order (thr falls as the list fills), one round each among the
    // keys above the last taken (keys are unique); with FAST2 the two smallest go in first
    auto insert = [&](uint32_t mk) __attribute__((always_inline)) {
        if (KO) {
            list_insert_u32<KL>(Lk, mk);
        } else {
            const uint32_t ir = (uint32_t)t + (mk & 31u);
#pragma unroll
            for (int j = KL - 1; j > 0; --j) {
                const bool up = mk < Lk[j - 1];
                const bool here = mk < Lk[j];
                Li[j] = up ? Li[j - 1] : (here ? ir : Li[j]);
                Lk[j] = max(Lk[j - 1], min(mk, Lk[j]));
            }
            if (mk < Lk[0]) {
                Li[0] = ir;
                Lk[0] = mk;
            }
        }
        thr = min(thr, Lk[KL - 1]);
    };
    uint32_t mn = m1;
    if (FAST2) {
        if (m1 < thr) insert(m1);
        if (!__any(m2 < thr)) return;
        if (m2 < thr) insert(m2);
        mn = kKey32None;
#pragma unroll
        for (int r = 0; r < 16; ++r) mn = min(mn, key[r] > m2 ? key[r] : kKey32None);
    }
    while (__any(mn < thr)) {
        if (mn < thr) insert(mn);
        const uint32_t last = mn;
        mn = kKey32None;
#pragma unroll
        for (int r = 0; r < 16; ++r) mn = min(mn, key[r] > last ? key[r] : kKey32None);
    }
}

// Half-wave merge: lanes l and l+32 hold lists of the same query; lane l writes the KL smallest
// of both to out[j] (the query's list for this slice, contiguous). Keys a list dropped are >= its
// last key >= the merged last key, so
// the merged last key still bounds every dropped candidate from below. With gq (the query's shared
// bound, see share_issue) a full merged list publishes its last key there.
template <int KL, bool KO = false>
__device__ __forceinline__ void half_merge_out(uint64_t (*lds)[KL], int half, const uint32_t (&Lk)[KL],
                                               const uint32_t (&Li)[KL], uint64_t *out, uint32_t *gq = nullptr,
                                               int64_t rb = 0, uint32_t lowm = 31u) {
    // list entry e of this lane as a 64-bit list key (approximation word, row)
    auto entry = [&](int e) __attribute__((always_inline)) {
        const uint32_t row = KO ? (uint32_t)rb + (Lk[e] & lowm) : Li[e];
        return Lk[e] == kKey32None ? kKey64None : (((uint64_t)Lk[e] << 32) | row);
    };
    if (half == 1) {
#pragma unroll
        for (int j = 0; j < KL; ++j) lds[0][j] = entry(j);
    }
    __syncthreads();
    if (half == 0) {
        int pb = 0;
        int pa = 0;
        for (int j = 0; j < KL; ++j) {
            uint64_t va = kKey64None;
#pragma unroll
            for (int e = 0; e < KL; ++e) // own list entry pa, selected without dynamic register indexing
                if (e == pa) va = entry(e);
            const uint64_t vb = pb < KL ? lds[0][pb] : kKey64None;
            const bool takea = va <= vb;
            out[j] = takea ? va : vb;
            pa += takea ? 1 : 0;
            pb += takea ? 0 : 1;
            if (gq && j == KL - 1 && (takea ? va : vb) != kKey64None) atomicMin(gq, (uint32_t)((takea ? va : vb) >> 32));
        }
    }
}

// Bound shared between the slices of one query: gthr[q] holds the smallest list bound any slice of
// query q has reached (the KL-th key of a full list of the query's rows, so all of the query's KL
// best keys lie below it). Every kShareTiles tiles a slice publishes its own bound there and takes
// the smaller of the two: a slice then skips the rows that the query's other slices have already
// beaten instead of filling its list with them. Dropped rows stay bounded by what the merge sees:
// a published bound is at or above the final last key of the publishing slice's full list. The
// atomic is issued before the tile's loads (its result is waited for after the products without
// draining them) and only by the first half-wave; both halves take the result.
constexpr int kShareTiles = 4;

// Row bits of keys-only lists: the mask of the smallest power of two holding rows_per_slice rows
// (at least the 5 bits of a tile offset).
__device__ __forceinline__ uint32_t slice_row_mask(int64_t rows_per_slice) {
    const uint32_t bits = 64 - __clzll((unsigned long long)max<int64_t>(rows_per_slice - 1, 31));
    return (1u << bits) - 1u;
}

__device__ __forceinline__ uint32_t share_issue(uint32_t *gthr, int64_t q, int half, uint32_t thr) {
    return half == 0 ? atomicMin(gthr + q, thr) : kKey32None;
}

__device__ __forceinline__ uint32_t share_take(uint32_t thr, uint32_t old) {
    return min(thr, min(old, (uint32_t)__shfl_xor((int)old, 32, 64)));
}

template <int KL, int PREC, bool KO = false>
__device__ __forceinline__ void scan_generic(uint64_t (*lds)[32][KL], const __bf16 *__restrict__ Xh,
                                             const __bf16 *__restrict__ Xl, const float *__restrict__ xn, int64_t n,
                                             int dp, const __bf16 *__restrict__ Qh, const __bf16 *__restrict__ Ql,
                                             const float *__restrict__ qn, int64_t rows_per_slice,
                                             uint64_t *__restrict__ part, int64_t qsl, int64_t qb, int64_t s,
                                             uint32_t *__restrict__ gthr = nullptr) {
    // A workgroup owns 128 queries (wave w: queries 32w..32w+31 of the block) and one slice of
    // rows; the four waves walk the same 32-row tiles, so each tile is fetched from L2 once per
    // workgroup and served to the other waves from L1. Row arrays are padded with zero rows to a
    // multiple of 32 and slices start on 32-row boundaries: tile loads need no guard, rows >= r1
    // are masked.
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, half = lane >> 5, col = lane & 31;
    const int64_t q = qb * 128 + 32 * w + col; // query buffers are padded to whole 128-query blocks
    const int64_t r0 = s * rows_per_slice, r1 = min(n, r0 + rows_per_slice);
    const uint32_t lowm = KO ? slice_row_mask(rows_per_slice) : 31u;
    const float qnv = qn[q];
    const __bf16 *qh = Qh + q * dp + 8 * half;
    const __bf16 *ql = Ql + q * dp + 8 * half;

    uint32_t Lk[KL], Li[KL];
#pragma unroll
    for (int j = 0; j < KL; ++j) {
        Lk[j] = kKey32None;
        Li[j] = 0u;
    }
    uint32_t thr = kKey32None;

    for (int64_t t = r0; t < r1; t += 32) {
        const bool share = gthr && (((t - r0) >> 5) % kShareTiles) == 0;
        uint32_t old = kKey32None;
        if (share) old = share_issue(gthr, q, half, thr);
        // |x|^2 of this lane's 16 output rows: rows t + 8g + 4 half + {0..3}
        float4 xv[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) xv[g] = *(const float4 *)(xn + t + 8 * g + 4 * half);
        const __bf16 *xh = Xh + (t + col) * dp + 8 * half;
        const __bf16 *xl = Xl + (t + col) * dp + 8 * half;
        f32x16 acc = {};
        mfma_tile_bf16<PREC, PREC == 3 ? 4 : 8>(acc, xh, xl, qh, ql, dp);
        thr = min(thr, (uint32_t)__shfl_xor((int)thr, 32, 64)); // the other half-wave's list of this query
        if (share) thr = share_take(thr, old);
        if (t == r0) // only a slice's last tile can hold rows past its end
            t + 32 <= r1 ? tile_keys_insert<KL, false, true, false, KO>(acc, xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm)
                         : tile_keys_insert<KL, true, true, false, KO>(acc, xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm);
        else
            t + 32 <= r1 ? tile_keys_insert<KL, false, false, false, KO>(acc, xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm)
                         : tile_keys_insert<KL, true, false, false, KO>(acc, xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm);
    }
    half_merge_out<KL, KO>(lds[w] + col, half, Lk, Li, part + q * qsl + s * KL, gthr ? gthr + q : nullptr, r0, lowm);
}

// The same scan for dp = 16 NS (NS <= 4): the lane's query fragments stay in registers for the
// whole slice, and row fragments and norms are loaded two tiles ahead into a ring of three
// register sets that rotate by name (three tiles per loop trip; a copy between sets would wait
// for the loads in flight), so the memory latency hides behind the work of the two tiles before.
template <int KL, int PREC, int NS, bool KO = false>
__device__ __forceinline__ void scan_small(uint64_t (*lds)[32][KL], const __bf16 *__restrict__ Xh,
                                           const __bf16 *__restrict__ Xl, const float *__restrict__ xn, int64_t n,
                                           const __bf16 *__restrict__ Qh, const __bf16 *__restrict__ Ql,
                                           const float *__restrict__ qn, int64_t rows_per_slice,
                                           uint64_t *__restrict__ part, int64_t qsl, int64_t qb, int64_t s,
                                           int g = 4, uint32_t *__restrict__ gthr = nullptr) {
    constexpr int dp = 16 * NS;
    constexpr int NPL = PREC == 3 ? 2 : 1;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, half = lane >> 5, col = lane & 31;
    // g query groups of 32 per workgroup (4; 2 or 1 for a block of at most 64 or 32 queries): wave w
    // takes group w % g of query block qb (of 32 g queries) and slice s (4 / g) + w / g
    const int64_t q = qb * 32 * g + 32 * (w % g) + col;
    s = s * (4 / g) + w / g;
    const int64_t r0 = s * rows_per_slice, r1 = min(n, r0 + rows_per_slice);
    const uint32_t lowm = KO ? slice_row_mask(rows_per_slice) : 31u;
    const float qnv = qn[q];
    bf16x8 b[NS][NPL];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        b[k][0] = *(const bf16x8 *)(Qh + q * dp + 16 * k + 8 * half);
        if (PREC == 3) b[k][NPL - 1] = *(const bf16x8 *)(Ql + q * dp + 16 * k + 8 * half);
    }
    uint32_t Lk[KL], Li[KL];
#pragma unroll
    for (int j = 0; j < KL; ++j) {
        Lk[j] = kKey32None;
        Li[j] = 0u;
    }
    uint32_t thr = kKey32None;
    // tiles i = 0 .. nt - 1 at rows r0 + 32 i; loads past the last tile re-read it (padded arrays:
    // in bounds), and up to two tiles past the end are computed on those rows and masked whole
    struct Tile {
        bf16x8 a[NS][NPL];
        float4 xv[4];
    };
    // addresses: a wave-uniform tile base plus a loop-invariant 32-bit lane offset (scalar base,
    // vector offset and immediate loads, no 64-bit arithmetic per load)
    const int aoff = col * dp + 8 * half, noff = 4 * half;
    auto load_tile = [&](int64_t t, Tile &f) __attribute__((always_inline)) {
        const __bf16 *th = Xh + t * dp, *tl = Xl + t * dp;
        const float *tn = xn + t;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            f.a[k][0] = *(const bf16x8 *)(th + aoff + 16 * k);
            if (PREC == 3) f.a[k][NPL - 1] = *(const bf16x8 *)(tl + aoff + 16 * k);
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) f.xv[g] = *(const float4 *)(tn + noff + 8 * g);
    };
    const int64_t nt = r0 < r1 ? (r1 - r0 + 31) / 32 : 0;
    auto step = [&](int64_t i, const Tile &cur, Tile &ahead) __attribute__((always_inline)) {
        const bool share = gthr && i % kShareTiles == 0;
        uint32_t old = kKey32None;
        if (share) old = share_issue(gthr, q, half, thr);
        load_tile(r0 + 32 * min(i + 2, nt - 1), ahead);
        const int64_t t = r0 + 32 * i;
        f32x16 acc = {};
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cur.a[k][0], b[k][0], acc, 0, 0, 0);
            if (PREC == 3) {
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cur.a[k][0], b[k][NPL - 1], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cur.a[k][NPL - 1], b[k][0], acc, 0, 0, 0);
            }
        }
        thr = min(thr, (uint32_t)__shfl_xor((int)thr, 32, 64)); // the other half-wave's list of this query
        if (share) thr = share_take(thr, old);
        if (t == r0) // only a slice's last tile can hold rows past its end
            t + 32 <= r1 ? tile_keys_insert<KL, false, true, false, KO>(acc, cur.xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm)
                         : tile_keys_insert<KL, true, true, false, KO>(acc, cur.xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm);
        else
            t + 32 <= r1 ? tile_keys_insert<KL, false, false, false, KO>(acc, cur.xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm)
                         : tile_keys_insert<KL, true, false, false, KO>(acc, cur.xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm);
    };
    if (nt > 0) {
        Tile f0, f1, f2;
        load_tile(r0, f0);
        load_tile(r0 + 32 * min((int64_t)1, nt - 1), f1);
        for (int64_t i = 0; i < nt; i += 3) {
            step(i, f0, f2);
            step(i + 1, f1, f0);
            step(i + 2, f2, f1);
        }
    }
    half_merge_out<KL, KO>(lds[w] + col, half, Lk, Li, part + q * qsl + s * KL, gthr ? gthr + q : nullptr, r0, lowm);
}

// The wide scan without LDS staging, for workgroups whose waves scan different slices (blocks of
// at most 64 queries: g = 1 or 2 query groups, as in scan_small): the row fragments stream
// through NC register sets, one per 64-column chunk of a tile, set c refilled with the next
// tile's chunk c right after its products.
template <int KL, int PREC, int NC, bool KO = false>
__device__ __forceinline__ void scan_wide_direct(uint64_t (*lds)[32][KL], const __bf16 *__restrict__ Xh,
                                                 const __bf16 *__restrict__ Xl, const float *__restrict__ xn,
                                                 int64_t n, const __bf16 *__restrict__ Qh,
                                                 const __bf16 *__restrict__ Ql, const float *__restrict__ qn,
                                                 int64_t rows_per_slice, uint64_t *__restrict__ part, int64_t qsl,
                                                 int64_t qb, int64_t s, int g, uint32_t *__restrict__ gthr = nullptr) {
    constexpr int dp = 64 * NC;
    constexpr int NPL = PREC == 3 ? 2 : 1;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, half = lane >> 5, col = lane & 31;
    const int64_t q = qb * 32 * g + 32 * (w % g) + col;
    s = s * (4 / g) + w / g;
    const int64_t r0 = s * rows_per_slice, r1 = min(n, r0 + rows_per_slice);
    const uint32_t lowm = KO ? slice_row_mask(rows_per_slice) : 31u;
    const float qnv = qn[q];
    bf16x8 b[4 * NC][NPL];
#pragma unroll
    for (int k = 0; k < 4 * NC; ++k) {
        b[k][0] = *(const bf16x8 *)(Qh + q * dp + 16 * k + 8 * half);
        if (PREC == 3) b[k][NPL - 1] = *(const bf16x8 *)(Ql + q * dp + 16 * k + 8 * half);
    }
    uint32_t Lk[KL], Li[KL];
#pragma unroll
    for (int j = 0; j < KL; ++j) {
        Lk[j] = kKey32None;
        Li[j] = 0u;
    }
    uint32_t thr = kKey32None;
    const int aoff = col * dp + 8 * half, noff = 4 * half;
    const int64_t nt = r0 < r1 ? (r1 - r0 + 31) / 32 : 0;
    if (nt > 0) {
        bf16x8 a[4 * NC][NPL]; // chunk c: a[4c .. 4c + 3]
#pragma unroll
        for (int k = 0; k < 4 * NC; ++k) {
            a[k][0] = *(const bf16x8 *)(Xh + r0 * dp + aoff + 16 * k);
            if (PREC == 3) a[k][NPL - 1] = *(const bf16x8 *)(Xl + r0 * dp + aoff + 16 * k);
        }
        float4 xv[4];
#pragma unroll
        for (int gi = 0; gi < 4; ++gi) xv[gi] = *(const float4 *)(xn + r0 + noff + 8 * gi);
        for (int64_t i = 0; i < nt; ++i) {
            const int64_t t = r0 + 32 * i;
            const int64_t tn = r0 + 32 * min(i + 1, nt - 1); // past the last tile: that tile again
            const __bf16 *th = Xh + tn * dp, *tl = Xl + tn * dp;
            const bool share = gthr && i % kShareTiles == 0;
            uint32_t old = kKey32None;
            if (share) old = share_issue(gthr, q, half, thr);
            f32x16 acc = {};
#pragma unroll
            for (int c = 0; c < NC; ++c) {
#pragma unroll
                for (int k = 4 * c; k < 4 * c + 4; ++k) {
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[k][0], b[k][0], acc, 0, 0, 0);
                    if (PREC == 3) {
                        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[k][0], b[k][NPL - 1], acc, 0, 0, 0);
                        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[k][NPL - 1], b[k][0], acc, 0, 0, 0);
                    }
                }
#pragma unroll
                for (int k = 4 * c; k < 4 * c + 4; ++k) {
                    a[k][0] = *(const bf16x8 *)(th + aoff + 16 * k);
                    if (PREC == 3) a[k][NPL - 1] = *(const bf16x8 *)(tl + aoff + 16 * k);
                }
            }
            thr = min(thr, (uint32_t)__shfl_xor((int)thr, 32, 64)); // the other half-wave's list of this query
            if (share) thr = share_take(thr, old);
            if (t == r0)
                t + 32 <= r1 ? tile_keys_insert<KL, false, true, false, KO>(acc, xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm)
                             : tile_keys_insert<KL, true, true, false, KO>(acc, xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm);
            else
                t + 32 <= r1 ? tile_keys_insert<KL, false, false, false, KO>(acc, xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm)
                             : tile_keys_insert<KL, true, false, false, KO>(acc, xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm);
#pragma unroll
            for (int gi = 0; gi < 4; ++gi) xv[gi] = *(const float4 *)(xn + tn + noff + 8 * gi);
        }
    }
    half_merge_out<KL, KO>(lds[w] + col, half, Lk, Li, part + q * qsl + s * KL, gthr ? gthr + q : nullptr, r0, lowm);
}

// LDS staging area of scan_wide (bf16 elements): two buffers of NPL planes of 32 rows of
// 64 NC + 8 columns (16 bytes of pad per row).
template <int NC, int NPL> constexpr int wide_stage_elems() { return 2 * NPL * 32 * (64 * NC + 8); }

// The same scan for dp = 64 NC (NC = 2..4). The lane's query fragments stay in registers for the
// whole slice. A tile's rows are staged once per workgroup into LDS by coalesced loads (16-byte
// pieces, consecutive threads along a row: whole 64-byte L1 accesses) and every wave reads its
// row fragments from there: direct fragment loads would touch 32 half-used segments per load and
// fetch each tile once per wave, and L1 accesses, not products, would bound the scan. The next
// tile's pieces are loaded into registers during this tile's products and stored into the other
// buffer after them (one barrier per tile; rows are padded: the loads need no guard).
template <int KL, int PREC, int NC, bool KO = false>
__device__ __forceinline__ void scan_wide(uint64_t (*lds)[32][KL], __bf16 *stage, const __bf16 *__restrict__ Xh,
                                          const __bf16 *__restrict__ Xl, const float *__restrict__ xn, int64_t n,
                                          const __bf16 *__restrict__ Qh, const __bf16 *__restrict__ Ql,
                                          const float *__restrict__ qn, int64_t rows_per_slice,
                                          uint64_t *__restrict__ part, int64_t qsl, int64_t qb, int64_t s,
                                          uint32_t *__restrict__ gthr = nullptr) {
    constexpr int dp = 64 * NC;
    constexpr int NPL = PREC == 3 ? 2 : 1;
    constexpr int RS = dp + 8;          // LDS row stride (bf16)
    constexpr int PLANE = 32 * RS, BUF = NPL * PLANE;
    constexpr int PPR = dp / 8;         // 16-byte pieces per row
    constexpr int NPC = NPL * NC;       // pieces per thread per tile (NPL x 32 x PPR / 256)
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, half = lane >> 5, col = lane & 31;
    const int64_t q = qb * 128 + 32 * w + col;
    const int64_t r0 = s * rows_per_slice, r1 = min(n, r0 + rows_per_slice);
    const uint32_t lowm = KO ? slice_row_mask(rows_per_slice) : 31u;
    const float qnv = qn[q];
    bf16x8 b[4 * NC][NPL];
#pragma unroll
    for (int k = 0; k < 4 * NC; ++k) {
        b[k][0] = *(const bf16x8 *)(Qh + q * dp + 16 * k + 8 * half);
        if (PREC == 3) b[k][NPL - 1] = *(const bf16x8 *)(Ql + q * dp + 16 * k + 8 * half);
    }
    uint32_t Lk[KL], Li[KL];
#pragma unroll
    for (int j = 0; j < KL; ++j) {
        Lk[j] = kKey32None;
        Li[j] = 0u;
    }
    uint32_t thr = kKey32None;
    const int64_t nt = r0 < r1 ? (r1 - r0 + 31) / 32 : 0;
    if (nt > 0) {
        // staging role: piece tid + 256 u is plane u / NC, then row and piece of that plane
        int goff[NPC], soff[NPC];
#pragma unroll
        for (int u = 0; u < NPC; ++u) {
            const int rem = tid + 256 * (u % NC), row = rem / PPR, pc = rem % PPR;
            goff[u] = row * dp + 8 * pc;
            soff[u] = (u / NC) * PLANE + row * RS + 8 * pc;
        }
        bf16x8 sreg[NPC];
        auto gload = [&](int64_t t) __attribute__((always_inline)) {
#pragma unroll
            for (int u = 0; u < NPC; ++u)
                sreg[u] = *(const bf16x8 *)((u / NC == 0 ? Xh : Xl) + t * dp + goff[u]);
        };
        auto sstore = [&](int bf) __attribute__((always_inline)) {
#pragma unroll
            for (int u = 0; u < NPC; ++u) *(bf16x8 *)(stage + bf * BUF + soff[u]) = sreg[u];
        };
        gload(r0);
        sstore(0);
        float4 xv[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) xv[g] = *(const float4 *)(xn + r0 + 8 * g + 4 * half);
        __syncthreads();
        for (int64_t i = 0; i < nt; ++i) {
            const int bf = (int)(i & 1);
            const int64_t t = r0 + 32 * i;
            const int64_t tn = r0 + 32 * min(i + 1, nt - 1);
            const bool share = gthr && i % kShareTiles == 0;
            uint32_t old = kKey32None;
            if (share) old = share_issue(gthr, q, half, thr);
            if (i + 1 < nt) gload(tn); // (in flight during the products)
            const __bf16 *sa = stage + bf * BUF + col * RS + 8 * half;
            f32x16 acc = {};
#pragma unroll
            for (int k = 0; k < 4 * NC; ++k) {
                const bf16x8 ah = *(const bf16x8 *)(sa + 16 * k);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, b[k][0], acc, 0, 0, 0);
                if (PREC == 3) {
                    const bf16x8 al = *(const bf16x8 *)(sa + PLANE + 16 * k);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, b[k][NPL - 1], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, b[k][0], acc, 0, 0, 0);
                }
            }
            thr = min(thr, (uint32_t)__shfl_xor((int)thr, 32, 64)); // the other half-wave's list of this query
            if (share) thr = share_take(thr, old);
            if (t == r0)
                t + 32 <= r1 ? tile_keys_insert<KL, false, true, false, KO>(acc, xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm)
                             : tile_keys_insert<KL, true, true, false, KO>(acc, xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm);
            else
                t + 32 <= r1 ? tile_keys_insert<KL, false, false, true, KO>(acc, xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm)
                             : tile_keys_insert<KL, true, false, true, KO>(acc, xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm);
#pragma unroll
            for (int g = 0; g < 4; ++g) xv[g] = *(const float4 *)(xn + tn + 8 * g + 4 * half);
            // the other buffer was last read for tile i - 1, which every wave finished before the
            // previous barrier
            if (i + 1 < nt) sstore(bf ^ 1);
            __syncthreads();
        }
    }
    half_merge_out<KL, KO>(lds[w] + col, half, Lk, Li, part + q * qsl + s * KL, gthr ? gthr + q : nullptr, r0, lowm);
}

// Keys-only tile keys without branches (the scan overlaps them with the next tile's products):
// the lane's keys of the tile (rows tb + offset in the slice), its three smallest, and the two
// smallest inserted where below thr (inserting the empty key changes nothing). A lane with its
// third smallest still below thr has more keys to take: tile_keys_rest.
template <int KL>
__device__ __forceinline__ void tile_keys_fast(const f32x16 &acc, const float4 (&xv)[4], float qnv, uint32_t tb,
                                               uint32_t lowm, int half, uint32_t (&Lk)[KL], uint32_t &thr,
                                               uint32_t (&key)[16], uint32_t &m2, uint32_t &m3) {
    uint32_t m1 = kKey32None;
    m2 = kKey32None;
    m3 = kKey32None;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const float xr = (r & 3) == 0 ? xv[r >> 2].x : (r & 3) == 1 ? xv[r >> 2].y : (r & 3) == 2 ? xv[r >> 2].z : xv[r >> 2].w;
        const float v = fmaxf(0.f, (qnv + xr) - 2.f * acc[r]);
        const int off = 8 * (r >> 2) + 4 * half + (r & 3);
        const uint32_t k = (f32_key(v) & ~lowm) | (tb + (uint32_t)off);
        key[r] = k;
        m3 = min(m3, max(m2, k));
        m2 = min(m2, max(m1, k));
        m1 = min(m1, k);
    }
    list_insert_u32<KL>(Lk, m1 < thr ? m1 : kKey32None);
    thr = min(thr, Lk[KL - 1]);
    list_insert_u32<KL>(Lk, m2 < thr ? m2 : kKey32None);
    thr = min(thr, Lk[KL - 1]);
}

// The keys above m2 still below thr (keys are unique): by a sort and merge when a lane holds
// several (lists of 16 or 32), else one round per key.
template <int KL>
__device__ __forceinline__ void tile_keys_rest(uint32_t (&key)[16], uint32_t m2, uint32_t (&Lk)[KL], uint32_t &thr) {
    if constexpr (KL == 16 || KL == 32) {
        int below = 0;
#pragma unroll
        for (int r = 0; r < 16; ++r) below += key[r] > m2 && key[r] < thr ? 1 : 0;
        if (__any(below >= 3)) {
#pragma unroll
            for (int r = 0; r < 16; ++r) key[r] = key[r] > m2 && key[r] < thr ? key[r] : kKey32None;
            sort16(key);
            uint32_t unused[KL];
            merge_sorted16<KL, true>(key, 0u, Lk, unused);
            thr = min(thr, Lk[KL - 1]);
            return;
        }
    }
    uint32_t last = m2;
    for (;;) {
        uint32_t mn = kKey32None;
#pragma unroll
        for (int r = 0; r < 16; ++r) mn = min(mn, key[r] > last ? key[r] : kKey32None);
        if (!__any(mn < thr)) break;
        if (mn < thr) {
            list_insert_u32<KL>(Lk, mn);
            thr = min(thr, Lk[KL - 1]);
        }
        last = mn;
    }
}

// scan_wide for keys-only lists, software-pipelined: a tile's keys are taken in the same
// straight-line block as the next tile's products, so the key arithmetic issues while the matrix
// cores work (a wave per SIMD: there is no other wave to fill them), and only lanes with a third
// key below their bound leave it. Tiles are staged as in scan_wide, one tile further ahead; full
// tiles go through the pipeline, a slice's partial last tile (rows past n) after it with masking.
template <int KL, int PREC, int NC>
__device__ __forceinline__ void scan_wide_pipe(uint64_t (*lds)[32][KL], __bf16 *stage, const __bf16 *__restrict__ Xh,
                                               const __bf16 *__restrict__ Xl, const float *__restrict__ xn, int64_t n,
                                               const __bf16 *__restrict__ Qh, const __bf16 *__restrict__ Ql,
                                               const float *__restrict__ qn, int64_t rows_per_slice,
                                               uint64_t *__restrict__ part, int64_t qsl, int64_t qb, int64_t s,
                                               uint32_t *__restrict__ gthr = nullptr) {
    constexpr int dp = 64 * NC;
    constexpr int NPL = PREC == 3 ? 2 : 1;
    constexpr int RS = dp + 8;
    constexpr int PLANE = 32 * RS, BUF = NPL * PLANE;
    constexpr int PPR = dp / 8;
    constexpr int NPC = NPL * NC;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, half = lane >> 5, col = lane & 31;
    const int64_t q = qb * 128 + 32 * w + col;
    const int64_t r0 = s * rows_per_slice, r1 = min(n, r0 + rows_per_slice);
    const uint32_t lowm = slice_row_mask(rows_per_slice);
    const float qnv = qn[q];
    bf16x8 b[4 * NC][NPL];
#pragma unroll
    for (int k = 0; k < 4 * NC; ++k) {
        b[k][0] = *(const bf16x8 *)(Qh + q * dp + 16 * k + 8 * half);
        if (PREC == 3) b[k][NPL - 1] = *(const bf16x8 *)(Ql + q * dp + 16 * k + 8 * half);
    }
    uint32_t Lk[KL], Li[KL]; // (Li: unused, keys-only lists)
#pragma unroll
    for (int j = 0; j < KL; ++j) Lk[j] = kKey32None;
    uint32_t thr = kKey32None;
    const int64_t nt = r0 < r1 ? (r1 - r0 + 31) / 32 : 0; // tiles, the last one possibly partial
    const int64_t nf = r0 < r1 ? (r1 - r0) / 32 : 0;      // full tiles
    if (nt > 0) {
        int goff[NPC], soff[NPC];
#pragma unroll
        for (int u = 0; u < NPC; ++u) {
            const int rem = tid + 256 * (u % NC), row = rem / PPR, pc = rem % PPR;
            goff[u] = row * dp + 8 * pc;
            soff[u] = (u / NC) * PLANE + row * RS + 8 * pc;
        }
        bf16x8 sreg[NPC];
        auto gload = [&](int64_t t) __attribute__((always_inline)) {
#pragma unroll
            for (int u = 0; u < NPC; ++u)
                sreg[u] = *(const bf16x8 *)((u / NC == 0 ? Xh : Xl) + t * dp + goff[u]);
        };
        auto sstore = [&](int bf) __attribute__((always_inline)) {
#pragma unroll
            for (int u = 0; u < NPC; ++u) *(bf16x8 *)(stage + bf * BUF + soff[u]) = sreg[u];
        };
        auto products = [&](int bf) __attribute__((always_inline)) {
            const __bf16 *sa = stage + bf * BUF + col * RS + 8 * half;
            f32x16 acc = {};
#pragma unroll
            for (int k = 0; k < 4 * NC; ++k) {
                const bf16x8 ah = *(const bf16x8 *)(sa + 16 * k);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, b[k][0], acc, 0, 0, 0);
                if (PREC == 3) {
                    const bf16x8 al = *(const bf16x8 *)(sa + PLANE + 16 * k);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, b[k][NPL - 1], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, b[k][0], acc, 0, 0, 0);
                }
            }
            return acc;
        };
        auto norms = [&](int64_t t, float4 (&xv)[4]) __attribute__((always_inline)) {
#pragma unroll
            for (int g = 0; g < 4; ++g) xv[g] = *(const float4 *)(xn + t + 8 * g + 4 * half);
        };
        // prologue: tile 0 in buffer 0 and its products; tile 1 in buffer 1
        gload(r0);
        sstore(0);
        if (nt > 1) gload(r0 + 32);
        float4 xv[4];
        norms(r0, xv);
        __syncthreads();
        f32x16 acc = products(0);
        if (nt > 1) sstore(1);
        __syncthreads();
        for (int64_t i = 0; i < nf; ++i) {
            const int64_t t = r0 + 32 * i;
            const bool share = gthr && i % kShareTiles == 0;
            uint32_t old = kKey32None;
            if (share) old = share_issue(gthr, q, half, thr);
            if (i + 2 < nt) gload(r0 + 32 * (i + 2)); // (in flight during this step)
            float4 xn1[4];
            norms(r0 + 32 * min(i + 1, nt - 1), xn1);
            // the next tile's products (past the last tile: the stale buffer, unused) and this
            // tile's keys, one block
            const f32x16 an = products((int)((i + 1) & 1));
            thr = min(thr, (uint32_t)__shfl_xor((int)thr, 32, 64)); // the other half-wave's list of this query
            thr = share_take(thr, old); // (no branch: old is the empty key when not shared)
            uint32_t key[16], m2, m3;
            tile_keys_fast<KL>(acc, xv, qnv, (uint32_t)(t - r0), lowm, half, Lk, thr, key, m2, m3);
            // issue order of the block: each product followed by an LDS read and a few of the key
            // operations, so that those issue while the matrix core works
#pragma unroll
            for (int k = 0; k < 4 * NC * (PREC == 3 ? 3 : 1); ++k) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0); // one MFMA
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0); // one LDS read
                __builtin_amdgcn_sched_group_barrier(0x002, 5, 0); // five VALU
            }
            if (__any(m3 < thr)) tile_keys_rest<KL>(key, m2, Lk, thr);
            acc = an;
#pragma unroll
            for (int g = 0; g < 4; ++g) xv[g] = xn1[g];
            // tile i + 2 into the buffer tile i was read from (every wave took its products in
            // the previous step, before the barrier that ended it)
            if (i + 2 < nt) sstore((int)(i & 1));
            __syncthreads();
        }
        if (nt > nf) { // the partial last tile (its products are in acc): rows past r1 masked
            const int64_t t = r0 + 32 * nf;
            thr = min(thr, (uint32_t)__shfl_xor((int)thr, 32, 64));
            nf == 0 ? tile_keys_insert<KL, true, true, false, true>(acc, xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm)
                    : tile_keys_insert<KL, true, false, false, true>(acc, xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm);
        }
    }
    half_merge_out<KL, true>(lds[w] + col, half, Lk, Li, part + q * qsl + s * KL, gthr ? gthr + q : nullptr, r0, lowm);
}

// scan_small for keys-only lists, software-pipelined as scan_wide_pipe: a tile's keys are taken
// in the straight-line block of the next tile's products. The row fragments and norms come from
// a ring of three register sets loaded two tiles ahead (a set holds tile i for its norms, i + 1 for
// the products, and i + 2 is loaded into the set of i - 1); g query groups per workgroup as in
// scan_small.
template <int KL, int PREC, int NS>
__device__ __forceinline__ void scan_small_pipe(uint64_t (*lds)[32][KL], const __bf16 *__restrict__ Xh,
                                                const __bf16 *__restrict__ Xl, const float *__restrict__ xn, int64_t n,
                                                const __bf16 *__restrict__ Qh, const __bf16 *__restrict__ Ql,
                                                const float *__restrict__ qn, int64_t rows_per_slice,
                                                uint64_t *__restrict__ part, int64_t qsl, int64_t qb, int64_t s,
                                                int g = 4, uint32_t *__restrict__ gthr = nullptr) {
    constexpr int dp = 16 * NS;
    constexpr int NPL = PREC == 3 ? 2 : 1;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, half = lane >> 5, col = lane & 31;
    const int64_t q = qb * 32 * g + 32 * (w % g) + col;
    s = s * (4 / g) + w / g;
    const int64_t r0 = s * rows_per_slice, r1 = min(n, r0 + rows_per_slice);
    const uint32_t lowm = slice_row_mask(rows_per_slice);
    const float qnv = qn[q];
    bf16x8 b[NS][NPL];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        b[k][0] = *(const bf16x8 *)(Qh + q * dp + 16 * k + 8 * half);
        if (PREC == 3) b[k][NPL - 1] = *(const bf16x8 *)(Ql + q * dp + 16 * k + 8 * half);
    }
    uint32_t Lk[KL], Li[KL]; // (Li: unused, keys-only lists)
#pragma unroll
    for (int j = 0; j < KL; ++j) Lk[j] = kKey32None;
    uint32_t thr = kKey32None;
    struct Tile {
        bf16x8 a[NS][NPL];
        float4 xv[4];
    };
    const int aoff = col * dp + 8 * half, noff = 4 * half;
    const int64_t nt = r0 < r1 ? (r1 - r0 + 31) / 32 : 0; // tiles, the last one possibly partial
    const int64_t nf = r0 < r1 ? (r1 - r0) / 32 : 0;      // full tiles
    auto load_tile = [&](int64_t i, Tile &f) __attribute__((always_inline)) {
        const int64_t t = r0 + 32 * min(i, nt - 1); // (past the last tile: that tile again, in bounds)
        const __bf16 *th = Xh + t * dp, *tl = Xl + t * dp;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            f.a[k][0] = *(const bf16x8 *)(th + aoff + 16 * k);
            if (PREC == 3) f.a[k][NPL - 1] = *(const bf16x8 *)(tl + aoff + 16 * k);
        }
#pragma unroll
        for (int gg = 0; gg < 4; ++gg) f.xv[gg] = *(const float4 *)(xn + t + noff + 8 * gg);
    };
    auto products = [&](const Tile &f) __attribute__((always_inline)) {
        f32x16 acc = {};
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.a[k][0], b[k][0], acc, 0, 0, 0);
            if (PREC == 3) {
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.a[k][0], b[k][NPL - 1], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.a[k][NPL - 1], b[k][0], acc, 0, 0, 0);
            }
        }
        return acc;
    };
    if (nt > 0) {
        Tile f0, f1, f2;
        load_tile(0, f0);
        load_tile(1, f1);
        f32x16 acc = products(f0);
        // step i: tile i's keys (acc, its norms in `cur`), tile i + 1's products (`nxt`), tile i + 2
        // loaded into `ld` (the set of tile i - 1)
        auto step = [&](int64_t i, const Tile &cur, const Tile &nxt, Tile &ld) __attribute__((always_inline)) {
            const bool share = gthr && i % kShareTiles == 0;
            uint32_t old = kKey32None;
            if (share) old = share_issue(gthr, q, half, thr);
            load_tile(i + 2, ld);
            const f32x16 an = products(nxt);
            thr = min(thr, (uint32_t)__shfl_xor((int)thr, 32, 64)); // the other half-wave's list of this query
            thr = share_take(thr, old); // (no branch: old is the empty key when not shared)
            uint32_t key[16], m2, m3;
            tile_keys_fast<KL>(acc, cur.xv, qnv, (uint32_t)(32 * i), lowm, half, Lk, thr, key, m2, m3);
            if (__any(m3 < thr)) tile_keys_rest<KL>(key, m2, Lk, thr);
            acc = an;
        };
        int64_t i = 0;
        for (; i + 3 <= nf; i += 3) {
            step(i, f0, f1, f2);
            step(i + 1, f1, f2, f0);
            step(i + 2, f2, f0, f1);
        }
        // the last full tiles (at most two) and the partial one, whose products are then in acc
        // and its norms in the set that the rotation reached
        if (i < nf) {
            step(i, f0, f1, f2);
            ++i;
            if (i < nf) {
                step(i, f1, f2, f0);
                ++i;
                if (nt > nf) {
                    const int64_t t = r0 + 32 * nf;
                    thr = min(thr, (uint32_t)__shfl_xor((int)thr, 32, 64));
                    tile_keys_insert<KL, true, false, false, true>(acc, f2.xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm);
                }
            } else if (nt > nf) {
                const int64_t t = r0 + 32 * nf;
                thr = min(thr, (uint32_t)__shfl_xor((int)thr, 32, 64));
                tile_keys_insert<KL, true, false, false, true>(acc, f1.xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm);
            }
        } else if (nt > nf) {
            const int64_t t = r0 + 32 * nf;
            thr = min(thr, (uint32_t)__shfl_xor((int)thr, 32, 64));
            nf == 0 ? tile_keys_insert<KL, true, true, false, true>(acc, f0.xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm)
                    : tile_keys_insert<KL, true, false, false, true>(acc, f0.xv, qnv, t, r1, half, Lk, Li, thr, r0, lowm);
        }
    }
    half_merge_out<KL, true>(lds[w] + col, half, Lk, Li, part + q * qsl + s * KL, gthr ? gthr + q : nullptr, r0, lowm);
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}

// Exact squared sums (the reference's FMA chains, as l2_core_group8) of the candidates in lanes
// 0 .. nkeep-1 for rows of d in [8 ST, 8 ST + 7]: the query's chain elements stay in registers and
// each 8-lane group runs two candidates a round (16 a round for the wave) with all row loads
// issued before either chain. A missing candidate reads row 0 (n >= 1) and its sum is dropped.
// Candidate j's sum ends in lane j.
template <int ST>
__device__ __forceinline__ float rerank_cores_st(const float *__restrict__ qrow, const float *__restrict__ X,
                                                 int64_t xstride, int64_t d, int64_t n, int32_t myidx, int nkeep,
                                                 int lane) {
    const int grp = lane >> 3, c8 = lane & 7;
    float qa[ST];
#pragma unroll
    for (int u = 0; u < ST; ++u) qa[u] = qrow[c8 + 8 * u];
    float core = 0.f;
    for (int base = 0; base < nkeep; base += 16) {
        const int ca = base + grp, cb = base + 8 + grp;
        const int32_t ia = __shfl(myidx, ca & 63, 64), ib = __shfl(myidx, cb & 63, 64);
        const float *xa = X + (ca < nkeep && ia >= 0 && ia < n ? (int64_t)ia : 0) * xstride;
        const float *xb = X + (cb < nkeep && ib >= 0 && ib < n ? (int64_t)ib : 0) * xstride;
        float ta[ST], tb[ST];
#pragma unroll
        for (int u = 0; u < ST; ++u) {
            ta[u] = xa[c8 + 8 * u];
            tb[u] = xb[c8 + 8 * u];
        }
        float aa = 0.f, ab = 0.f;
#pragma unroll
        for (int u = 0; u < ST; ++u) {
            const float t = qa[u] - ta[u], s = qa[u] - tb[u];
            aa = fmaf(t, t, aa);
            ab = fmaf(s, s, ab);
        }
        const float va = l2_group8_finish(aa, qrow, xa, d, c8);
        const float vb = l2_group8_finish(ab, qrow, xb, d, c8);
        const int j = lane - base; // candidate ca's owner for j in 0..7, cb's for 8..15
        const float ga = __shfl(va, (j & 7) * 8, 64), gb = __shfl(vb, (j & 7) * 8, 64);
        if (j >= 0 && j < 8) core = ga;
        else if (j >= 8 && j < 16) core = gb;
    }
    return core;
}

// rerank_cores_st for rows of a multiple of 256 floats: the chains run through the row in chunks of
// 32 steps (the query's elements of a chunk re-read from cache, the two candidates' loaded
// together), so 16 candidates a round fit in registers at any such width.
__device__ __forceinline__ float rerank_cores_chunked(const float *__restrict__ qrow, const float *__restrict__ X,
                                                      int64_t xstride, int64_t d, int64_t n, int32_t myidx,
                                                      int nkeep, int lane) {
    constexpr int CH = 32;
    const int grp = lane >> 3, c8 = lane & 7;
    float core = 0.f;
    for (int base = 0; base < nkeep; base += 16) {
        const int ca = base + grp, cb = base + 8 + grp;
        const int32_t ia = __shfl(myidx, ca & 63, 64), ib = __shfl(myidx, cb & 63, 64);
        const float *xa = X + (ca < nkeep && ia >= 0 && ia < n ? (int64_t)ia : 0) * xstride;
        const float *xb = X + (cb < nkeep && ib >= 0 && ib < n ? (int64_t)ib : 0) * xstride;
        float aa = 0.f, ab = 0.f;
        for (int64_t e0 = c8; e0 < d; e0 += 8 * CH) {
            float qa[CH], ta[CH], tb[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                qa[u] = qrow[e0 + 8 * u];
                ta[u] = xa[e0 + 8 * u];
                tb[u] = xb[e0 + 8 * u];
            }
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const float t = qa[u] - ta[u], s = qa[u] - tb[u];
                aa = fmaf(t, t, aa);
                ab = fmaf(s, s, ab);
            }
        }
        const float va = l2_group8_finish(aa, qrow, xa, d, c8);
        const float vb = l2_group8_finish(ab, qrow, xb, d, c8);
        const int j = lane - base;
        const float ga = __shfl(va, (j & 7) * 8, 64), gb = __shfl(vb, (j & 7) * 8, 64);
        if (j >= 0 && j < 8) core = ga;
        else if (j >= 8 && j < 16) core = gb;
    }
    return core;
}

// One wave per query q: exact rerank of K2c candidates, (distance, index) order, completeness check.
// REG: rows of 8 u + (0..7) floats for u in {4, 8, 16, 24, 32} go through rerank_cores_st; rows of
// a multiple of 256 floats (other widths) through rerank_cores_chunked.
// Candidate j (j < K2c) is the key `mine` of lane j: its approximation's bits in the high word, its
// row in the low word (empty key: no candidate).
template <bool REG>
__device__ __forceinline__ void rerank_regs(const float *__restrict__ X, int64_t xstride, int64_t d,
                                            const float *__restrict__ Q, int64_t qstride,
                                            const float *__restrict__ qn, const unsigned int *__restrict__ xn_max_bits,
                                            uint64_t mine, int K2c, const uint64_t *__restrict__ part, int64_t nlists,
                                            int KL, int64_t qsl, int kk, int64_t n, float eps_coef,
                                            int32_t *__restrict__ out_idx, float *__restrict__ out_dist,
                                            int32_t *__restrict__ flags, int64_t q, uint32_t lowm = 31u) {
    // lowm: the row bits of the keys (31: offsets in a tile; keys-only lists: offsets in a slice)
    const int lane = threadIdx.x & 63;
    // the smallest approximation any full list dropped: its last key with the row bits cleared
    uint32_t bmin = kKey32None;
    for (int64_t l = lane; l < nlists; l += 64) {
        const uint32_t last = (uint32_t)(part[q * qsl + l * KL + KL - 1] >> 32);
        bmin = min(bmin, last == kKey32None ? kKey32None : (last & ~lowm));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) bmin = min(bmin, (uint32_t)__shfl_xor((int)bmin, off, 64));

    // Candidates come sorted by approximation (a key's high word holds it with the low bits lowm
    // replaced by row bits, so the true approximation lies between the word with those bits clear
    // and with them set). Every approximation is within eps of the exact value, so a candidate
    // whose approximation exceeds the kk-th candidate's by more than 2 eps has an exact value
    // above each of the first kk candidates' and cannot be among the kk nearest: its row is not
    // read. The kept candidates are a prefix.
    const float sq = sqrtf(qn[q]) + sqrtf(__uint_as_float(*xn_max_bits));
    const float eps = eps_coef * sq * sq;
    int32_t myidx = lane < K2c && mine != kKey64None ? (int32_t)(uint32_t)(mine & 0xFFFFFFFFull) : -1;
    if (K2c > kk) {
        const uint32_t hk = (uint32_t)(__shfl(mine, kk - 1, 64) >> 32);
        if (hk != kKey32None && lane < K2c) {
            const float upper = key_f32(hk | lowm);
            const float lower = key_f32((uint32_t)(mine >> 32) & ~lowm);
            if (lower - upper > 3.f * eps) myidx = -1;
        }
    }
    const uint32_t al = (uint32_t)(__shfl(mine, K2c - 1, 64) >> 32); // the last candidate's approximation
    const uint64_t kept = __ballot(myidx >= 0);
    const int nkeep = kept ? 64 - __clzll((long long)kept) : 0;
    const float *qrow = Q + q * qstride;
    float core = 0.f;
    switch (REG ? d >> 3 : 0) {
    case 4: core = rerank_cores_st<4>(qrow, X, xstride, d, n, myidx, nkeep, lane); break;
    case 8: core = rerank_cores_st<8>(qrow, X, xstride, d, n, myidx, nkeep, lane); break;
    case 16: core = rerank_cores_st<16>(qrow, X, xstride, d, n, myidx, nkeep, lane); break;
    case 24: core = rerank_cores_st<24>(qrow, X, xstride, d, n, myidx, nkeep, lane); break;
    case 32: core = rerank_cores_st<32>(qrow, X, xstride, d, n, myidx, nkeep, lane); break;
    default: {
        if (d >= 256 && d % 256 == 0) {
            core = rerank_cores_chunked(qrow, X, xstride, d, n, myidx, nkeep, lane);
            break;
        }
        // 8-lane groups (lane c of a group runs the reference's FMA chain c, reading 8
        // consecutive floats with its group), 8 candidates a round; candidate j's sum ends in lane j
        const int grp = lane >> 3, c8 = lane & 7;
        for (int base = 0; base < nkeep; base += 8) {
            const int cand = base + grp;
            const int32_t idx = __shfl(myidx, cand & 63, 64);
            float v = 0.f;
            if (cand < K2c && idx >= 0 && idx < n) // uniform over the group
                v = l2_core_group8(qrow, X + (int64_t)idx * xstride, d, c8);
            const bool owner = lane >= base && lane < base + 8;
            const float got = __shfl(v, owner ? 8 * (lane - base) : 0, 64);
            if (owner) core = got;
        }
    }
    }
    uint64_t key = kKey64None;
    if (lane < K2c && myidx >= 0 && myidx < n) key = ((uint64_t)f32_key(sqrtf(core)) << 32) | (uint32_t)myidx;
    float dk_core = 0.f;
    if (kk >= 4) {
        // one bitonic sort of the wave's 64 keys: lane r holds the r-th smallest
        uint64_t v[1] = {key};
        wave_sort_u64<1, 2>(v, lane);
        if (lane < kk) {
            out_idx[q * kk + lane] = (int32_t)(uint32_t)(v[0] & 0xFFFFFFFFull);
            out_dist[q * kk + lane] = key_f32((uint32_t)(v[0] >> 32));
        }
        const uint64_t mk = __shfl(v[0], kk - 1, 64);
        const unsigned long long own = __ballot(key == mk);
        dk_core = __shfl(core, __ffsll(own) - 1, 64);
    } else {
        for (int r = 0; r < kk; ++r) {
            const uint64_t m = wave_min_u64(key);
            const unsigned long long own = __ballot(key == m);
            const int owner = __ffsll((unsigned long long)own) - 1;
            const float cr = __shfl(core, owner, 64);
            if (key == m) key = kKey64None;
            if (lane == 0) {
                out_idx[q * kk + r] = (int32_t)(uint32_t)(m & 0xFFFFFFFFull);
                out_dist[q * kk + r] = key_f32((uint32_t)(m >> 32));
            }
            dk_core = cr;
        }
    }
    if (lane == 0) {
        int flag = 0;
        if (n > K2c) {
            const uint32_t bnd = min(bmin, al == kKey32None ? kKey32None : (al & ~lowm));
            const float bound = bnd == kKey32None ? INFINITY : key_f32(bnd);
            // every row outside the candidates has core >= bound - eps; the output orders by
            // sqrtf(core), and a root equal to the k-th's would tie and could win on its lower
            // index, so the roots must differ as well (taken from just below bound - eps, which
            // absorbs the subtraction's rounding; the roots are monotone in their arguments)
            const float lo = fmaxf(bound - eps, 0.f);
            flag = !(lo > dk_core && sqrtf(nextafterf(lo, 0.f)) > sqrtf(dk_core));
        }
        flags[q] = flag;
    }
}

// rerank_regs over candidates stored as rows cidx[q, :K2c] (-1: none) and approximations
// capx[q, :K2c].
template <bool REG>
__device__ __forceinline__ void rerank_query(const float *__restrict__ X, int64_t xstride, int64_t d,
                                             const float *__restrict__ Q, int64_t qstride,
                                             const float *__restrict__ qn, const unsigned int *__restrict__ xn_max_bits,
                                             const int32_t *__restrict__ cidx, const float *__restrict__ capx, int K2c,
                                             const uint64_t *__restrict__ part, int64_t nlists, int KL, int64_t qsl,
                                             int kk, int64_t n, float eps_coef, int32_t *__restrict__ out_idx,
                                             float *__restrict__ out_dist, int32_t *__restrict__ flags, int64_t q,
                                             uint32_t lowm = 31u) {
    const int lane = threadIdx.x & 63;
    uint64_t mine = kKey64None;
    if (lane < K2c) {
        const int32_t ci = cidx[q * K2c + lane];
        if (ci >= 0) mine = ((uint64_t)f32_key(capx[q * K2c + lane]) << 32) | (uint32_t)ci;
    }
    rerank_regs<REG>(X, xstride, d, Q, qstride, qn, xn_max_bits, mine, K2c, part, nlists, KL, qsl, kk, n, eps_coef,
                     out_idx, out_dist, flags, q, lowm);
}

// Exact full scan of query q by one workgroup (called for flagged queries only).
template <int K>
__device__ __forceinline__ void fallback_query(uint64_t (*wres)[K], const float *__restrict__ X, int64_t xstride,
                                               int64_t n, int64_t d, const float *__restrict__ Q, int64_t qstride,
                                               int kk, int32_t *__restrict__ out_idx, float *__restrict__ out_dist,
                                               int64_t q) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t L[K];
#pragma unroll
    for (int j = 0; j < K; ++j) L[j] = kKey64None;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const float dd = sqrtf(l2_core(Q + q * qstride, X + i * xstride, d));
        const uint64_t key = ((uint64_t)f32_key(dd) << 32) | (uint32_t)i;
        if (key < L[K - 1]) list_insert_u64<K>(L, key);
    }
    for (int r = 0; r < kk; ++r) { // wave-level extraction of the wave's kk best
        const uint64_t m = wave_min_u64(L[0]);
        if (L[0] == m && m != kKey64None) {
#pragma unroll
            for (int j = 0; j < K - 1; ++j) L[j] = L[j + 1];
            L[K - 1] = kKey64None;
        }
        if (lane == 0) wres[w][r] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int p[4] = {0, 0, 0, 0};
        auto head = [&](int ww) { return p[ww] < kk ? wres[ww][p[ww]] : kKey64None; };
        for (int r = 0; r < kk; ++r) {
            int bw = 0;
            for (int ww = 1; ww < 4; ++ww)
                if (head(ww) < head(bw)) bw = ww;
            const uint64_t m = head(bw);
            ++p[bw];
            out_idx[q * kk + r] = (int32_t)(uint32_t)(m & 0xFFFFFFFFull);
            out_dist[q * kk + r] = key_f32((uint32_t)(m >> 32));
        }
    }
    __syncthreads(); // wres is rewritten by the workgroup's next query
}

// ---------------------------------------------------------------- host side
} // namespace pn
