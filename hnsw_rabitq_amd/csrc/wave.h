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

// dist_batch's 8-bit path split in two for c <= 32 ids, so that the row loads can be issued well
// before the distances are needed: the issue step starts the loads of a lane's two slots (grp and
// grp + 16; an inactive slot reads row 0 and its result is dropped), dist_u8_lanes computes.
// The ids are on lanes (lane 32 + s holds the id of slot s, s < c): each lane fetches its two
// slots' ids by permutes.
template <int NCH>
__device__ __forceinline__ void dist_u8_issue_lanes(const DevGraph& g, uint32_t ids_on_lanes, int c, int lane,
                                                    uint32_t (&w0)[NCH], uint32_t (&w1)[NCH]) {
    const int grp = lane >> 2, sub = lane & 3;
    const uint32_t a = (uint32_t)__shfl((int)ids_on_lanes, 32 + grp),
                   b = (uint32_t)__shfl((int)ids_on_lanes, 48 + grp);
    const uint32_t id0 = grp < c ? a : 0u, id1 = grp + 16 < c ? b : 0u;
    load_u8_words<NCH>(g.vec8 + (size_t)id0 * (16 * NCH) + sub * 4 * NCH, w0);
    load_u8_words<NCH>(g.vec8 + (size_t)id1 * (16 * NCH) + sub * 4 * NCH, w1);
}

// Exact integer form of an 8-bit row's distance to a query whose coordinates are integers in
// 0..255 as well (qw: the lane's query words in the row layout, sq = sum of the query's squares).
// Every step of the FP32 computation is then exact: a difference is an integer of magnitude at
// most 255, its square at most 65,025, and every partial sum of at most 16 * NCH <= 256 such
// terms stays below 2^24, where FP32 holds integers exactly (the same for the products and sums
// of the inner product).  So the FP32 distance equals the integer sum(q - x)^2 = sum q^2 +
// sum x^2 - 2 sum q.x (L2), resp. 1 - sum q.x (IP) with that integer converted exactly, and four
// byte products per instruction give it bit for bit.
template <int SPACE, int NCH>
__device__ __forceinline__ float dist_group16_u8int(const uint32_t (&w)[NCH], const uint32_t (&qw)[NCH], int sq) {
    static_assert(16 * NCH <= 256, "partial sums must stay below 2^24");
    uint32_t qx = 0u, xx = 0u;
#pragma unroll
    for (int t = 0; t < NCH; t++) {
        qx = __builtin_amdgcn_udot4(qw[t], w[t], qx, false);
        if (SPACE == 0) xx = __builtin_amdgcn_udot4(w[t], w[t], xx, false);
    }
    int v = SPACE == 0 ? (int)xx - 2 * (int)qx : (int)qx;
    // the four lanes of the group (quad permutations [1,0,3,2] and [2,3,0,1])
    v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
    return SPACE == 0 ? (float)(v + sq) : 1.0f - (float)v;
}

// The query's words in the 8-bit row layout for this lane (coordinates 16t+4g..16t+4g+3 of
// lane g = lane & 3) and sum of its squares; returns whether every coordinate of the query is
// an integer in 0..255 with exactly that FP32 value (uniform over the wave).
template <int NCH>
__device__ __forceinline__ bool query_u8_words(const float* qs, int lane, uint32_t (&qw)[NCH], int& sq) {
    const int g = lane & 3;
    bool ok = true;
#pragma unroll
    for (int t = 0; t < NCH; t++) {
        const float4 q = *reinterpret_cast<const float4*>(qs + 16 * t + 4 * g);
        const float qv[4] = {q.x, q.y, q.z, q.w};
        uint32_t w = 0u;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const bool in = qv[j] >= 0.f && qv[j] <= 255.f;
            const uint32_t b = in ? (uint32_t)qv[j] : 0u;
            ok &= in && __float_as_uint((float)b) == __float_as_uint(qv[j]);
            w |= b << (8 * j);
        }
        qw[t] = w;
    }
    uint32_t s = 0u;
#pragma unroll
    for (int t = 0; t < NCH; t++) s = __builtin_amdgcn_udot4(qw[t], qw[t], s, false);
    int v = (int)s;
    v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
    sq = v;
    return __builtin_amdgcn_ballot_w64(!ok) == 0ull;
}

// The distances of slots grp and grp + 16 (every lane of a group holds its group's two), moved so
// that lane c returns the distance of slot c & 31 (lanes c and 32 + c the same).
template <int SPACE, int NCH>
__device__ __forceinline__ float dist_u8_lanes(const float* qs, const uint32_t (&w0)[NCH], const uint32_t (&w1)[NCH],
                                               int lane, bool qint, const uint32_t (&qw)[NCH], int sq) {
    const int sub = lane & 3;
    float r0, r1;
    if (qint) {
        r0 = dist_group16_u8int<SPACE, NCH>(w0, qw, sq);
        r1 = dist_group16_u8int<SPACE, NCH>(w1, qw, sq);
    } else {
        r0 = dist_group16_u8<SPACE, NCH>(w0, qs, sub, lane);
        r1 = dist_group16_u8<SPACE, NCH>(w1, qs, sub, lane);
    }
    const int src = 4 * (lane & 15);
    const float a = __shfl(r0, src), b = __shfl(r1, src);
    return (lane & 16) ? b : a;
}

// The FP32 rows' counterpart for c <= 32 ids on lanes (lane 32 + s holds the id of slot s): loads
// and distances of the lane's two slots at once; lane c returns the distance of slot c & 31.
template <int SPACE, int NCH>
__device__ __forceinline__ float dist_f32_lanes(const DevGraph& g, const float* qs, uint32_t ids_on_lanes, int c,
                                                int lane) {
    const int grp = lane >> 2, sub = lane & 3;
    const uint32_t a = (uint32_t)__shfl((int)ids_on_lanes, 32 + grp),
                   b = (uint32_t)__shfl((int)ids_on_lanes, 48 + grp);
    const uint32_t id0 = grp < c ? a : 0u, id1 = grp + 16 < c ? b : 0u;
    const float4* x0 = reinterpret_cast<const float4*>(g.vec + (size_t)id0 * g.stride);
    const float4* x1 = reinterpret_cast<const float4*>(g.vec + (size_t)id1 * g.stride);
    float4 v0[NCH], v1[NCH];
#pragma unroll
    for (int t = 0; t < NCH; t++) v0[t] = x0[4 * t + sub];
#pragma unroll
    for (int t = 0; t < NCH; t++) v1[t] = x1[4 * t + sub];
    __builtin_amdgcn_sched_barrier(0);
    const float r0 = dist_group16_regs<SPACE, NCH>(v0, qs, sub, lane);
    const float r1 = dist_group16_regs<SPACE, NCH>(v1, qs, sub, lane);
    const int src = 4 * (lane & 15);
    const float ra = __shfl(r0, src), rb = __shfl(r1, src);
    return (lane & 16) ? rb : ra;
}

// ------------------------------------------------------------------------------------------
// Wave helpers
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ int lanes_below(uint64_t m, int lane) {
    return __popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
}

// Lane L's 64-bit value (L uniform).
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int L) {
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(v >> 32), L) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, L);
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

// Segment heads of a beam: the beam is cut into segments of 2^sh entries (2^sh * 64 >= the most
// entries it can hold), and lane c keeps the flag-free key of entry c << sh (~0 past the end), so
// that locating a key among the segments takes register compares instead of LDS probes.
__device__ __forceinline__ uint64_t beam_head(const uint64_t* bk, int n, int sh, int lane) {
    const int i = lane << sh;
    const uint64_t v = bk[i < n ? i : 0];
    return i < n ? (v & ~BFLAGS) : ~0ull;
}

// Number of beam entries whose key is below the lane's key, for every lane of `lanes` (other
// lanes return 0): the segment from the heads (one ballot per key of `lanes`, in registers), then
// a fixed-step search of sh steps inside it.  Equals beam_lower_bound64.
__device__ __forceinline__ int beam_lower_bound_heads(const uint64_t* bk, int n, int sh, uint64_t head,
                                                      uint64_t key, uint64_t lanes, int lane) {
    int seg = 0;  // heads below the key: entry (seg - 1) << sh is the last head below it
    const uint32_t klo = (uint32_t)key, khi = (uint32_t)(key >> 32);
    for (uint64_t t = lanes; t; t &= t - 1) {
        const int L = __ffsll((unsigned long long)t) - 1;
        const uint64_t kl = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)khi, L) << 32) |
                            (uint32_t)__builtin_amdgcn_readlane((int)klo, L);
        const int c = __popcll(__builtin_amdgcn_ballot_w64(head < kl));
        seg = lane == L ? c : seg;
    }
    // bk[base] < key; the count lies in (base, base + 2^sh]
    const int base = (seg > 0 ? seg - 1 : 0) << sh;
    const int lim = base + (1 << sh) < n ? base + (1 << sh) : n;
    int q = base;
    for (int step = (1 << sh) >> 1; step > 0; step >>= 1) {
        const int probe = q + step;
        const int pc = probe < lim ? probe : base;
        const bool lt = (probe < lim) & ((bk[pc] & ~BFLAGS) < key);
        q += lt ? step : 0;
    }
    return seg > 0 ? q + 1 : 0;
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
