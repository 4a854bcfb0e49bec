// Wave-level helpers shared by the search and build kernels (gfx950, 64-lane wavefronts).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"
#include "dist.h"

namespace hg {

// (da, ia) < (db, ib) in (distance, id) order; evaluated without branches.
__device__ __forceinline__ bool key_less(float da, uint32_t ia, float db, uint32_t ib) {
    return (int)(da < db) | ((int)(da == db) & (int)(ia < ib));
}

// One lane's NCH words of an 8-bit coded row (DevGraph::vec8): one contiguous run of 4*NCH bytes,
// read with the widest loads its alignment allows.
template <int NCH>
__device__ __forceinline__ void load_u8_words(const uint8_t* run, uint32_t (&w)[NCH]) {
    if constexpr (NCH % 4 == 0) {
#pragma unroll
        for (int t = 0; t < NCH; t += 4) {
            const uint4 v = reinterpret_cast<const uint4*>(run)[t / 4];
            w[t] = v.x;
            w[t + 1] = v.y;
            w[t + 2] = v.z;
            w[t + 3] = v.w;
        }
    } else if constexpr (NCH % 2 == 0) {
#pragma unroll
        for (int t = 0; t < NCH; t += 2) {
            const uint2 v = reinterpret_cast<const uint2*>(run)[t / 2];
            w[t] = v.x;
            w[t + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int t = 0; t < NCH; t++) w[t] = reinterpret_cast<const uint32_t*>(run)[t];
    }
}

// Distances of the c vectors whose ids are ids[0..c) to the query in LDS; results to out[].
// MODE 0 = generic (one lane per vector), otherwise dim % 16 == 0 with NCH chunks (0 = runtime).
// U8 (NCH > 0 only): the vectors are read from the 8-bit coded copy g.vec8.
template <int SPACE, int MODE, int NCH, bool U8 = false>
__device__ __forceinline__ void dist_batch(const DevGraph& g, const float* qs, const uint32_t* ids, int c,
                                           float* out, int lane) {
    static_assert(!U8 || (MODE == 1 && NCH > 0), "the 8-bit rows are read with a compile-time chunk count");
    if (MODE == 0) {
        for (int base = 0; base < c; base += 64) {
            int idx = base + lane;
            if (idx < c) {
                const float* x = g.vec + (size_t)ids[idx] * g.stride;
                out[idx] = dist_generic(SPACE, qs, x, (int)g.dim);
            }
        }
        return;
    }
    const int grp = lane >> 2, sub = lane & 3;
    const int nch = (int)(g.dim >> 4);
    for (int base = 0; base < c; base += 32) {
        // two vectors per group per round: slots grp and grp + 16
        int i0 = base + grp, i1 = base + grp + 16;
        bool a0 = i0 < c, a1 = i1 < c;
        uint32_t id0 = a0 ? ids[i0] : 0u, id1 = a1 ? ids[i1] : 0u;
        float r0, r1;
        if constexpr (U8) {
            // a lane's run of the coded row: 4*NCH bytes at (sub * NCH) words; an inactive slot
            // reads row 0 and its result is dropped
            constexpr int RUN = 4 * NCH;
            const uint8_t* x0 = g.vec8 + (size_t)id0 * (16 * NCH) + sub * RUN;
            const uint8_t* x1 = g.vec8 + (size_t)id1 * (16 * NCH) + sub * RUN;
            uint32_t w0[NCH], w1[NCH];
            load_u8_words<NCH>(x0, w0);
            load_u8_words<NCH>(x1, w1);
            __builtin_amdgcn_sched_barrier(0);
            r0 = dist_group16_u8<SPACE, NCH>(w0, qs, sub, lane);
            r1 = dist_group16_u8<SPACE, NCH>(w1, qs, sub, lane);
        } else if constexpr (NCH > 0) {
            // both vectors' chunks are loaded before either is used, with no branch around a
            // load: an inactive slot reads row 0 and its result is dropped
            const float4* x0 = reinterpret_cast<const float4*>(g.vec + (size_t)id0 * g.stride);
            const float4* x1 = reinterpret_cast<const float4*>(g.vec + (size_t)id1 * g.stride);
            float4 v0[NCH], v1[NCH];
#pragma unroll
            for (int t = 0; t < NCH; t++) v0[t] = x0[4 * t + sub];
#pragma unroll
            for (int t = 0; t < NCH; t++) v1[t] = x1[4 * t + sub];
            __builtin_amdgcn_sched_barrier(0);
            r0 = dist_group16_regs<SPACE, NCH>(v0, qs, sub, lane);
            r1 = dist_group16_regs<SPACE, NCH>(v1, qs, sub, lane);
        } else {
            dist_group16<SPACE, NCH>(g.vec + (size_t)id0 * g.stride, qs, nch, a0, sub, r0, lane);
            dist_group16<SPACE, NCH>(g.vec + (size_t)id1 * g.stride, qs, nch, a1, sub, r1, lane);
        }
        if (sub == 0) {
            if (a0) out[i0] = r0;
            if (a1) out[i1] = r1;
        }
    }
}

// ------------------------------------------------------------------------------------------
// Wave helpers
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ int lanes_below(uint64_t m, int lane) {
    return __popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
}

// Visited hash.  A forgotten node costs one re-evaluation (its key is then either in the beam,
// where the merge drops it, or no better than a bound that only tightens), never a different
// result.  PROBE = false: one slot per id, the id replaces whatever the slot held and counts as
// fresh unless the slot held it already (one LDS round trip).  PROBE = true: up to four probes
// from an empty-slot insert before an overwrite, so a table larger than the visited set forgets
// nothing (the distinct-evaluation count of the traversal).
template <bool PROBE = false>
__device__ __forceinline__ bool hash_insert(uint32_t* hash, uint32_t mask, uint32_t id) {
    const uint32_t h = (id * 2654435761u) & mask;
    if (!PROBE) return atomicExch(&hash[h], id) != id;
#pragma unroll 1
    for (int p = 0; p < 4; p++) {
        const uint32_t slot = (h + p) & mask;
        const uint32_t old = atomicCAS(&hash[slot], SENT, id);
        if (old == SENT) return true;
        if (old == id) return false;
    }
    hash[h] = id;  // window full: overwrite
    return true;
}

// Beam entries of the query kernel: one 64-bit key per entry ordered as (distance, id): the
// distance's order-preserving bits in the high word, id << 2 and two flag bits in the low word
// (BX: expanded, BD: not in the result set).  Keys of distinct nodes compare as (distance, id)
// whatever their flags (the flags sit below the id); the keys of one node differ only in the
// flags.  One 8-byte LDS access and one 64-bit compare per entry.
constexpr uint64_t BX = 1ull, BD = 2ull, BFLAGS = 3ull;
// float -> uint32 with the same order (-0 taken as +0; the distances are never NaN)
__device__ __forceinline__ uint32_t f2ord(float f) {
    const uint32_t u = __float_as_uint(f + 0.0f);
    return u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t k) {
    return __uint_as_float(k ^ ((k & 0x80000000u) ? 0x80000000u : 0xFFFFFFFFu));
}
__device__ __forceinline__ uint64_t bkey(float d, uint32_t id) { return ((uint64_t)f2ord(d) << 32) | ((uint64_t)id << 2); }
__device__ __forceinline__ float bdist(uint64_t k) { return ord2f((uint32_t)(k >> 32)); }
__device__ __forceinline__ uint32_t bid(uint64_t k) { return (uint32_t)k >> 2; }

// Number of beam entries whose key (flags aside) is below `key` (a flag-free key), over
// beam[0, n): a fixed-step lower bound (the probe index is clamped, so every lane runs the same
// steps with no branch).
__device__ __forceinline__ int beam_lower_bound64(const uint64_t* bk, int n, uint64_t key) {
    int pos = 0;
    for (int step = 1 << (31 - __clz(n > 0 ? n : 1)); step > 0; step >>= 1) {
        const int probe = pos + step - 1;
        const int pc = probe < n ? probe : (n > 0 ? n - 1 : 0);
        const bool lt = (probe < n) & ((bk[pc] & ~BFLAGS) < key);
        pos += lt ? step : 0;
    }
    return pos;
}

// Number of beam entries with key < (d, id), over beam[0, n): a fixed-step lower bound (the
// probe index is clamped, so every lane runs the same steps with no branch).
__device__ __forceinline__ int beam_lower_bound(const float* bd, const uint32_t* bi, int n, float d, uint32_t id) {
    int pos = 0;
    for (int step = 1 << (31 - __clz(n > 0 ? n : 1)); step > 0; step >>= 1) {
        const int probe = pos + step - 1;
        const int pc = probe < n ? probe : (n > 0 ? n - 1 : 0);
        const bool lt = (probe < n) & key_less(bd[pc], bi[pc] & ID_MASK, d, id);
        pos += lt ? step : 0;
    }
    return pos;
}

}  // namespace hg
