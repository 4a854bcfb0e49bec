// Batched HNSW query on gfx950: one wavefront per query, persistent workgroups.
//
// Restates HierarchicalNSW::searchKnn (hnswalg.h:1269-1323): entry distance, greedy descent
// through the upper layers (hnswalg.h:1277-1302, strict '<', first minimum in list order), then
// the level-0 best-first search searchBaseLayerST<bare_bone=true> (hnswalg.h:308-440) with
// ef_eff = max(ef, k), truncation to k and closest-first rows (bindings.cpp:649-654).
//
// Level-0 search state lives in LDS, per wave:
//   beam   : the reference's top_candidates as a list sorted by (distance, id), at most ef_eff
//            long; bit 31 of an id marks "expanded".  Under the (distance, id) total order the
//            reference's candidate_set is exactly the unexpanded part of this list (entries it
//            pops after they left top_candidates are beyond lowerBound and end the search), and
//            the next node it expands is the first unexpanded entry.
//   hash   : visited set (open addressing over node ids).  A lookup never reports a node that
//            was not inserted; when a probe window is full the slot is overwritten, so a node
//            may be evaluated twice.  That cannot change results: a re-evaluated node is either
//            still in the beam (dropped by the duplicate check of the merge) or was rejected or
//            evicted under a lowerBound that has only decreased since.
//   stage  : the neighbours of the expanded node that were not visited and their distances.
// A batch of neighbours is merged set-wise: under a total order, processing the neighbours
// one by one with the admission test `size < ef || d < lowerBound` and evicting the maximum
// leaves exactly the ef smallest keys of (beam U batch) -- what the merge computes.
//
// Distances follow the reference's FP32 order exactly (see dist_* below); all comparisons use
// (distance, id), so ties between equal distances resolve by node id.
#include <algorithm>

#include "common.h"
#include "dist.h"
#include "kernels.h"
#include "wave.h"

namespace hg {

// ------------------------------------------------------------------------------------------
// Learned stop: 12 traversal features and an in-wave MLP evaluation.
// ------------------------------------------------------------------------------------------
struct StopState {
    float qnorm, d_entry;
    int improved_recent;  // bitmask of the last 16 expansions that changed the top-k
    int last_change;      // expansion count at the last change of the top-k
    int win_pos;          // smallest beam position a new key took in the current window
    int prev_pos[3];      // the same for the three windows before
};

__device__ __forceinline__ float safe_div(float a, float b) { return a / (fabsf(b) + 1e-20f); }

// Features (same order as the trace collector and the host-side training code):
//  0 query norm                       1 log(1 + entry distance at level 0)
//  2 d_k / d_1                        3 d_ef / d_k (last beam entry)
//  4 hops / ef_eff                    5 fraction of the last 16 expansions that changed the top-k
//  6 d_next / d_k (next expansion)    7 position of the next expansion in the beam / ef_eff
//  8 d_k / d_entry
//  9 log(1 + smallest position a new key took in the current window / k)   (near misses)
// 10 the same over the last four windows
// 11 log(1 + number of beam entries closer than 1.2 d_k / k)                 (crowding at d_k)
// A window is the `period` expansions between two evaluations; a window without insertions
// counts as position ef_eff.
__device__ __forceinline__ float stop_feature(int f, const StopState& st, float d1, float dk, float dlast,
                                              float dnext, int pnext, int k, int hops, int ef_eff, int crowd) {
    const float fk = (float)k;
    const int w1 = st.win_pos < ef_eff ? st.win_pos : ef_eff;
    int w4 = w1;
    for (int j = 0; j < 3; j++) w4 = st.prev_pos[j] < w4 ? st.prev_pos[j] : w4;
    switch (f) {
        case 0: return st.qnorm;
        case 1: return logf(1.0f + st.d_entry);
        case 2: return safe_div(dk, d1);
        case 3: return safe_div(dlast, dk);
        case 4: return (float)hops / (float)ef_eff;
        case 5: return (float)__popc((uint32_t)st.improved_recent & 0xFFFFu) / 16.0f;
        case 6: return safe_div(dnext, dk);
        case 7: return (float)pnext / (float)ef_eff;
        case 8: return safe_div(dk, st.d_entry);
        case 9: return logf(1.0f + (float)w1 / fk);
        case 10: return logf(1.0f + (float)w4 / fk);
        default: return logf(1.0f + (float)crowd / fk);
    }
}

// The 12 features of the current beam bk[0, n) (sorted, every entry below scan_lo expanded) into
// feat[0..11].
__device__ __forceinline__ void beam_features(const uint64_t* bk, int n, int scan_lo, int kq, int ef_eff, int hops,
                                              const StopState& st, float* feat, int lane) {
    const int kk = kq < n ? kq : n;
    const float d1 = bdist(bk[0]), dk = bdist(bk[kk - 1]), dlast = bdist(bk[n - 1]);
    // the next expansion: first unexpanded entry (the search's stop rule ends the loop when there
    // is none)
    float dnext = dlast;
    int pnext = n;
    for (int b = scan_lo; b < n; b += 64) {
        const int i = b + lane;
        const uint64_t m = ballot(i < n && (bk[i] & BX) == 0);
        if (m) {
            pnext = b + __ffsll((unsigned long long)m) - 1;
            dnext = bdist(bk[pnext]);
            break;
        }
    }
    const int crowd = beam_lower_bound64(bk, n, bkey(1.2f * dk, 0u));
    if (lane < kStopFeatures) feat[lane] = stop_feature(lane, st, d1, dk, dlast, dnext, pnext, kq, hops, ef_eff, crowd);
}

__device__ __forceinline__ float gelu_tanh(float x) {
    const float c = 0.7978845608028654f;
    return 0.5f * x * (1.0f + tanhf(c * (x + 0.044715f * x * x * x)));
}

// FP32 -> the nearest BF16 value (ties to even), as FP32: the network's weights as the search
// kernel and the training kernels use them.
__device__ __forceinline__ float bf16_round(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xFFFF0000u);
}

// Logit of the network for the (normalised) features in LDS: lane j computes hidden unit j, its
// weight rows read from the FP32 master copy (cached; about 20 KB shared by every wave) and
// rounded to BF16 as they arrive, accumulation in FP32 -- the same operation order as the
// training forward pass (sequential FMAs over inputs, output accumulated in unit order).  The
// weights are not held in registers across the search: one evaluation every `period`
// expansions does not pay for the ~80 VGPRs (203 against 121 for the whole kernel, and two waves
// per SIMD fewer) that would take.  (A transposed copy read with coalesced loads took 144 VGPRs
// and was slower.)
__device__ __noinline__ float mlp_logit(const DevStopModel& m, const float* feat, float* hbuf, int lane) {
    const int H = m.hidden;
    const float* W1 = m.params;
    const float* b1 = W1 + H * kStopFeatures;
    const float* W2 = b1 + H;
    const float* b2 = W2 + H * H;
    const float* w3 = b2 + H;
    if (lane < H) {
        float s = b1[lane];
#pragma unroll
        for (int i = 0; i < kStopFeatures; i++) s = fma_(bf16_round(W1[lane * kStopFeatures + i]), feat[i], s);
        hbuf[lane] = gelu_tanh(s);
    }
    __syncthreads();
    float h2 = 0.f;
    if (lane < H) {
        float s = b2[lane];
        const float* row = W2 + lane * H;
#pragma unroll 8
        for (int i = 0; i < H; i++) s = fma_(bf16_round(row[i]), hbuf[i], s);
        h2 = gelu_tanh(s);
    }
    __syncthreads();
    if (lane < H) hbuf[lane] = h2;
    __syncthreads();
    const float w3l = lane < H ? bf16_round(w3[lane]) : 0.f;
    float o = 0.f;
    for (int j = 0; j < H; j++)
        o = fma_(hbuf[j], __int_as_float(__builtin_amdgcn_readlane(__float_as_int(w3l), j)), o);
    __syncthreads();
    return o + w3[H];
}

// ------------------------------------------------------------------------------------------
// Level-0 search of the plain and learned kernels (dim % 16 == 0 with a compiled chunk count,
// maxM0 <= 32): lazy merge, and on the 8-bit rows pipelined expansions.
//
// The beam is the sorted main list in LDS (bk[0, n)) together with up to 31 pending keys held in
// registers, one per lane, sorted over lanes 0..P-1 (pk: key with its expanded flag; pp: its
// insertion point in the main list).  A pending key's rank in the union is its lane plus its
// insertion point, so the union's ef-th key -- the reference's lowerBound once top_candidates is
// full -- is either a pending key of rank ef - 1 or the main entry at ef - 1 - (pending keys
// below it), which lies in the main list's last 64 entries (kept in registers, `tail`).  Keys of
// rank >= ef are those the reference evicted: pending ones are dropped at once, main ones are
// never expanded (the search stops at the first unexpanded key above the bound) and fall off at
// the next merge.  New keys join the pending list by one permute per 32-bit half; the main list is
// rewritten (the shift of the eager merge) only when 32 keys are pending and at the end, instead
// of after every expansion that admits a key.  Same admissions, same expansions, same result as
// the eager merge: the union at every step holds exactly the eager beam plus keys of rank >= ef.
//
// Pipelined expansions (U8): when the next expansion is known before this one ends (node2, the second
// unexpanded key, whose row lanes 32..63 loaded at the start of this expansion, with no new key
// before it), its neighbours are checked against the visited hash and their row loads issued at
// once (pre); the next expansion starts from them.  The visited hash is not touched by a merge,
// so checking them early finds the same fresh set.  The batch moves between lanes by permutes:
// lane 32 + s holds fresh slot s's id and then its distance.  A query of integers in 0..255 takes
// the exact integer form of the distances.
//
// STOPV (learned stop): every `period` expansions the pending keys are merged, so that the
// features are computed from the eager beam itself, and the network decides whether to stop.
// The union rank of a batch's smallest new key is its eager insertion position, which the
// near-miss and change features record.
// ------------------------------------------------------------------------------------------
template <int SPACE, int NCH, bool STOPV, bool U8>
__device__ __forceinline__ void level0_lazy(const DevGraph& g, const float* qs, uint64_t* bk, uint32_t* hash,
                                            uint32_t hmask, bool probe, uint64_t* trash_base, int ef_eff, int hcap,
                                            int lane, int& n_io, int& H0, int& E0, int& scan_lo_io,
                                            const SearchParams& Pm, StopState& st, int kq, float* feat,
                                            float* hbuf, bool& stopped) {
    auto visit = [&](uint32_t id) {
        return probe ? hash_insert<true>(hash, hmask, id) : hash_insert<false>(hash, hmask, id);
    };
    int n = n_io;
    int sh = 0;  // segments of 2^sh main entries, at most 32 of them
    while ((32 << sh) < ef_eff) sh++;
    uint64_t head = 0, tail = 0;  // head: lanes 32..63, the key of main entry (lane - 32) << sh
    int tb = 0;  // tail lane l holds main entry tb + l (flag-free; ~0 past the end)
    auto load_main_regs = [&]() {
        {
            const int i = (lane & 31) << sh;  // used on lanes 32..63 only
            const uint64_t v = bk[lane >= 32 && i < n ? i : 0];
            head = lane >= 32 && i < n ? (v & ~BFLAGS) : ~0ull;
        }
        tb = n > 64 ? n - 64 : 0;
        const int i = tb + lane;
        const uint64_t v = bk[i < n ? i : 0];
        tail = i < n ? (v & ~BFLAGS) : ~0ull;
    };
    load_main_regs();
    int P = 0;           // pending keys (lanes 0..P-1)
    uint64_t pk = ~0ull;  // pending key with its expanded flag
    int pp = 0;          // its insertion point in the main list
    uint64_t T = ~0ull;  // the union's ef-th key once the union holds ef keys, else ~0
    auto update_bound = [&]() {
        if (n + P < ef_eff) {
            T = ~0ull;
            return;
        }
        const int t = ef_eff - 1, U = lane + pp;
        const uint64_t eq = ballot(lane < P && U == t);
        if (eq) {
            T = readlane64(pk, __ffsll((unsigned long long)eq) - 1) & ~BFLAGS;
        } else {
            const int c = __popcll(ballot(lane < P && U < t));
            T = readlane64(tail, t - c - tb);
        }
    };
    update_bound();
    int scan_lo = scan_lo_io;
    // The scan window: 64 main entries from wb held one per lane (entries past the end read as
    // expanded), kept between merges -- the main list changes only in flush(), which drops it,
    // and in the expanded flags, which are set in the window too.
    int wb = -1;
    uint64_t we = 0;
    const int trash = (int)(trash_base - bk) + lane;
    // the pending keys into the main list (the eager merge's shift with the pending list as its
    // batch: lane i holds the key of rank i and its insertion point)
    auto flush = [&]() {
        if (P == 0) return;
        const int m = P, lim = ef_eff;
        const int new_n = n + m < lim ? n + m : lim;
        const int first_pos = __builtin_amdgcn_readlane(pp, 0);
        const int hi_chunk = (n - 1) & ~63;
        const int lo_chunk = first_pos & ~63;
        const uint64_t rank_lanes = (1ull << m) - 1ull;  // m <= 63
        const int spos = lane < m ? pp : 0x7FFFFFFF;
        for (int b0 = hi_chunk; b0 >= lo_chunk; b0 -= 8 * 64) {
            uint64_t vk[8];
            int mv[8];
            const int base = b0 + lane;
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int i = base - 64 * u;
                vk[u] = bk[i > 0 ? i : 0];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int cb = b0 - 64 * u;
                const int lo = __popcll(__builtin_amdgcn_ballot_w64(spos <= cb) & rank_lanes);
                const int hi = __popcll(__builtin_amdgcn_ballot_w64(spos <= cb + 63) & rank_lanes);
                int sv = lo;
                for (int rr = lo; rr < hi; rr++) sv += __builtin_amdgcn_readlane(spos, rr) <= base - 64 * u ? 1 : 0;
                mv[u] = sv;
            }
            __asm__ volatile("" ::: "memory");
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int i = base - 64 * u;
                const int np = i + mv[u];
                const bool w = (b0 - 64 * u >= lo_chunk) & (i >= first_pos) & (i < n) & (np < lim);
                bk[w ? np : trash] = vk[u];
            }
            __asm__ volatile("" ::: "memory");
        }
        const int fp = pp + lane;
        if (lane < m && fp < lim) bk[fp] = pk;
        n = new_n;
        scan_lo = first_pos < scan_lo ? first_pos : scan_lo;
        wb = -1;
        __syncthreads();
        load_main_regs();
        P = 0;
        pk = ~0ull;
        pp = 0;
        update_bound();
    };

    // pipelined expansions
    bool pre = false;
    int pre_cnt = 0;
    uint32_t pw0[U8 ? NCH : 1], pw1[U8 ? NCH : 1];
    uint32_t bid_c = SENT;  // lane 32 + s: the id of fresh slot s
    uint32_t pf_node = SENT, pf_row = SENT;  // lanes 32..63: the row of pf_node
    uint32_t qw[NCH];
    int qsq = 0;
    bool qint = false;
    if constexpr (U8) qint = query_u8_words<NCH>(qs, lane, qw, qsq);
    // the fresh ids of a row (lanes 0..31 hold its ids, SENT for none) onto lanes 32.. in list order
    auto stage_row = [&](uint32_t id) {
        const bool fresh = id != SENT && visit(id);
        const uint64_t fm = ballot(fresh);
        const int dst = fresh ? 32 + lanes_below(fm, lane) : (lane & 31);
        bid_c = (uint32_t)__builtin_amdgcn_ds_permute(dst * 4, (int)id);
        return __popcll(fm);
    };

    for (;;) {
        // the first unexpanded key of the main list (from scan_lo) and of the pending list, and the
        // ones after them; node2 = the second unexpanded of the union when it is known
        int pm = -1;
        uint64_t xm = ~0ull, xm2 = ~0ull;
        bool xm2_known = true;
        if (wb < 0 || scan_lo < wb || scan_lo >= wb + 64) wb = -1;
        for (int b = wb < 0 ? scan_lo : wb; b < n; b += 64) {
            if (b != wb) {
                wb = b;
                const int i = b + lane;
                we = i < n ? bk[i] : BX;
            }
            const uint64_t mu = ballot((b + lane >= scan_lo) & ((we & BX) == 0));
            if (mu) {
                const int l1 = __ffsll((unsigned long long)mu) - 1;
                pm = b + l1;
                xm = readlane64(we, l1) & ~BFLAGS;
                const uint64_t m2 = mu & (mu - 1);
                if (m2) xm2 = readlane64(we, __ffsll((unsigned long long)m2) - 1) & ~BFLAGS;
                xm2_known = m2 != 0 || b + 64 >= n;
                break;
            }
        }
        const uint64_t up = ballot(lane < P && (pk & BX) == 0);
        uint64_t xp = ~0ull, xp2 = ~0ull;
        int lp = -1;
        if (up) {
            lp = __ffsll((unsigned long long)up) - 1;
            xp = readlane64(pk, lp) & ~BFLAGS;
            const uint64_t u2 = up & (up - 1);
            if (u2) xp2 = readlane64(pk, __ffsll((unsigned long long)u2) - 1) & ~BFLAGS;
        }
        const bool from_main = xm < xp;
        const uint64_t x = from_main ? xm : xp;
        // no unexpanded key, or the first one is beyond the bound (evicted): the search ends
        if (x == ~0ull || x > T || H0 >= hcap) break;
        uint64_t k2;
        if (from_main) {
            if (lane == 0) bk[pm] = x | BX;
            if (lane == pm - wb) we |= BX;
            scan_lo = pm;
            k2 = xm2_known ? (xm2 < xp ? xm2 : xp) : ~0ull;
        } else {
            if (lane == lp) pk |= BX;
            k2 = xm < xp2 ? xm : xp2;
        }
        const uint32_t node = bid(x), node2 = k2 == ~0ull ? SENT : bid(k2);
        H0++;

        // neighbours: staged by the previous expansion, or from node's row (prefetched or loaded
        // now); lanes 32..63 load node2's row
        const bool staged = pre;
        pre = false;
        int cnt;
        {
            const int lr = lane & 31;
            const bool upper = lane >= 32;
            const uint32_t prev = (uint32_t)__shfl_xor((int)pf_row, 32);
            const bool need = upper ? node2 != SENT : (!staged && node != pf_node);
            uint32_t v = SENT;
            if (need && lr < (int)g.maxM0) v = g.links0[(size_t)(upper ? node2 : node) * g.maxM0 + lr];
            if (staged) {
                cnt = pre_cnt;
            } else {
                const uint32_t cur = node == pf_node ? prev : v;
                cnt = stage_row((!upper && lr < (int)g.maxM0) ? cur : SENT);
            }
            pf_row = v;
            pf_node = node2;
        }

        // distances, admission against the bound, insertion points in the main list, duplicates
        // (in the main list or pending), ranks; lanes 32..32+cnt-1 hold the batch
        int m = 0, cntb = 0, cp = 0, pos = 0, kmin_rank = 0;
        bool ok = false;
        uint64_t key = ~0ull, kmin = ~0ull;
        if (cnt > 0) {
            float dc;
            if constexpr (U8) {
                if (!staged) dist_u8_issue_lanes<NCH>(g, bid_c, cnt, lane, pw0, pw1);
                dc = dist_u8_lanes<SPACE, NCH>(qs, pw0, pw1, lane, qint, qw, qsq);
            } else {
                dc = dist_f32_lanes<SPACE, NCH>(g, qs, bid_c, cnt, lane);
            }
            E0 += cnt;
            if (lane >= 32 && lane - 32 < cnt) {
                key = bkey(dc, bid_c);
                if (!(key < T)) key = ~0ull;
            }
            const uint64_t live = ballot(key != ~0ull);
            if (live) {
                // per staged key, one ballot over the pending keys (lanes 0..31) and the segment
                // heads (lanes 32..63): the pending keys below it and its segment of the main list
                const uint64_t pkf = lane < P ? (pk & ~BFLAGS) : ~0ull;
                const uint64_t below = lane < 32 ? pkf : head;
                int sc = 0;  // (segment << 8) | pending keys below
                for (uint64_t t = live; t; t &= t - 1) {
                    const int L = __ffsll((unsigned long long)t) - 1;
                    const uint64_t b = __builtin_amdgcn_ballot_w64(below < readlane64(key, L));
                    const int v = (__popc((uint32_t)(b >> 32)) << 8) | __popc((uint32_t)b);
                    sc = lane == L ? v : sc;
                }
                const int seg = sc >> 8;
                cp = sc & 0xFF;
                // pending already: the pending key of rank cp equals it (lanes >= P hold ~0)
                const uint32_t plo = (uint32_t)__builtin_amdgcn_ds_bpermute(cp * 4, (int)(uint32_t)pkf);
                const uint32_t phi = (uint32_t)__builtin_amdgcn_ds_bpermute(cp * 4, (int)(uint32_t)(pkf >> 32));
                const bool dup = (((uint64_t)phi << 32) | plo) == key;
                // the count inside the segment: bk[base] < key, the count lies in (base, base + 2^sh];
                // found four ways per step (three probes read together), then one binary step
                // when sh is odd: 3 dependent LDS reads for segments of 32 or 64 entries
                const int base = (seg > 0 ? seg - 1 : 0) << sh;
                const int lim = base + (1 << sh) < n ? base + (1 << sh) : n;
                int q = base;
                int step = 1 << sh;
                while (step >= 4) {
                    step >>= 2;
                    const int p1 = q + step, p2 = q + 2 * step, p3 = q + 3 * step;
                    const uint64_t v1 = bk[p1 < lim ? p1 : base], v2 = bk[p2 < lim ? p2 : base],
                                   v3 = bk[p3 < lim ? p3 : base];
                    const int c = ((p1 < lim) & ((v1 & ~BFLAGS) < key)) + ((p2 < lim) & ((v2 & ~BFLAGS) < key)) +
                                  ((p3 < lim) & ((v3 & ~BFLAGS) < key));
                    q += c * step;
                }
                if (step == 2) {
                    const int p1 = q + 1;
                    q += ((p1 < lim) & ((bk[p1 < lim ? p1 : base] & ~BFLAGS) < key)) ? 1 : 0;
                }
                pos = seg > 0 ? q + 1 : 0;
                ok = key != ~0ull && !dup;
                if (ok && pos < n && (bk[pos] & ~BFLAGS) == key) ok = false;
                const uint64_t km = ballot(ok);
                m = __popcll(km);
                if (m > 0) {
                    // new-key lanes: rank among the new keys; pending lanes: new keys below them
                    const uint64_t vv = lane < 32 ? pkf : key;
                    for (uint64_t t = km; t; t &= t - 1) {
                        const int L = __ffsll((unsigned long long)t) - 1;
                        cntb += readlane64(key, L) < vv ? 1 : 0;
                    }
                    const int L0 = __ffsll((unsigned long long)ballot(ok && cntb == 0)) - 1;
                    kmin = readlane64(key, L0);
                    kmin_rank = __builtin_amdgcn_readlane(pos + cp, L0);
                }
            }
        }
        // the next expansion is node2 or the smallest new key, whichever is smaller; when it is
        // node2 (known, its row at hand) its neighbours are staged now, else its row is prefetched
        {
            const uint32_t next = kmin < k2 ? bid(kmin) : node2;
            // (the FP32 rows' loads would hold four times the registers across the merge: not staged)
            const bool stage = U8 && next != SENT && next == node2 && pf_node == node2;
            const uint32_t row = (uint32_t)__shfl_xor((int)pf_row, 32);
            uint32_t v = SENT;
            if (lane >= 32 && !stage && next != SENT && (lane & 31) < (int)g.maxM0)
                v = g.links0[(size_t)next * g.maxM0 + (lane & 31)];
            pf_row = v;
            pf_node = stage ? SENT : next;
            if (stage) {
                pre_cnt = stage_row((lane < 32 && lane < (int)g.maxM0) ? row : SENT);
                if constexpr (U8) {
                    if (pre_cnt > 0) dist_u8_issue_lanes<NCH>(g, bid_c, pre_cnt, lane, pw0, pw1);
                }
                pre = true;
            }
        }
        if (m > 0) {
            // the new keys join the pending list: pending lane i goes to i + (new keys below it),
            // a new key to (its rank among the new keys) + (pending keys below it); other lanes
            // park above
            const bool src_p = lane < P, src_n = lane >= 32 && ok;
            const uint64_t srcm = ballot(src_p || src_n);
            const int dst = src_p ? lane + cntb : (src_n ? cntb + cp : P + m + lanes_below(~srcm, lane));
            const uint64_t val = src_p ? pk : key;
            const int vpos = src_p ? pp : pos;
            const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_permute(dst * 4, (int)(uint32_t)val);
            const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_permute(dst * 4, (int)(uint32_t)(val >> 32));
            pp = __builtin_amdgcn_ds_permute(dst * 4, vpos);
            pk = ((uint64_t)hi << 32) | lo;
            P += m;
            // keys of union rank >= ef are evicted (a prefix of the pending list stays)
            P = __popcll(ballot(lane < P && lane + pp < ef_eff));
            if (lane >= P) pk = ~0ull;
            if constexpr (STOPV) {
                // the eager merge's record: the smallest new key's position and whether it
                // entered the top k of the beam it made
                const int live_n = n + P < ef_eff ? n + P : ef_eff;
                const int kk = kq < live_n ? kq : live_n;
                if (kmin_rank < kk) {
                    st.improved_recent |= 1;
                    st.last_change = H0;
                }
                st.win_pos = kmin_rank < st.win_pos ? kmin_rank : st.win_pos;
            }
            if (P >= 32)
                flush();
            else
                update_bound();
        }
        if constexpr (STOPV) {
            if (H0 % Pm.stop.period == 0) {
                // the network runs from min_hops on; before that only the windows advance
                const bool eval = H0 >= Pm.stop.min_hops;
                if (eval) {
                    flush();
                    beam_features(bk, n, scan_lo, kq, ef_eff, H0, st, feat, lane);
                }
                st.prev_pos[2] = st.prev_pos[1];
                st.prev_pos[1] = st.prev_pos[0];
                st.prev_pos[0] = st.win_pos < ef_eff ? st.win_pos : ef_eff;
                st.win_pos = 1 << 30;
                __syncthreads();
                if (eval) {
                    if (lane < kStopFeatures) feat[lane] = (feat[lane] - Pm.stop.mean[lane]) * Pm.stop.inv_std[lane];
                    __syncthreads();
                    const float logit = mlp_logit(Pm.stop, feat, hbuf, lane);
                    if (logit >= Pm.stop.logit_threshold) stopped = true;
                }
                __syncthreads();
            }
            st.improved_recent <<= 1;
            if (stopped) break;
        }
    }
    flush();
    n_io = n;
    scan_lo_io = scan_lo;
}

// Staged keys up to which the insertion points are located through the segment heads (a loop of
// one ballot per key) rather than by a full-depth LDS search per lane.
constexpr int kHeadKeys = 16;

// ------------------------------------------------------------------------------------------
// The search kernel.  Block = one wave (64 lanes) = one query at a time; workgroups loop over
// queries taken from a global counter until none are left.
// ------------------------------------------------------------------------------------------
// Membership of an internal id in the result set of the non-bare-bone search (hnswalg.h:325-326,
// 406-407): not marked deleted and allowed by the filter.
__device__ __forceinline__ bool id_allowed(uint32_t wdel, uint32_t wal, uint32_t id) {
    const uint32_t b = id & 31u;
    return ((wdel >> b) & 1u) == 0u && ((wal >> b) & 1u) != 0u;
}

// FILT: the non-bare-bone branch of searchBaseLayerST (hnswalg.h:323-437) in (distance, id)
// order.  Beam entries carry DFLAG when deleted or filtered out; such entries are traversed but
// never returned.  The beam ends right after the ef-th allowed entry once there are ef of them
// (nothing beyond it can be expanded: the reference stops at the first candidate farther than
// the result set's worst once that set is full); before that every admitted key is kept, up to
// the capacity `cap`; a query that needs more is abandoned and listed for a pass at a larger
// capacity (ovf_list), or, in the last pass, sets status bit 0.
template <int SPACE, int MODE, int NCH, bool STOP, bool TRACE, bool FILT, bool U8>
__global__ __launch_bounds__(64) void hnsw_search_kernel(SearchParams P) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const DevGraph& g = P.g;
    const uint32_t dimpad = (g.dim + 3u) & ~3u;
    const uint32_t cap = P.cap;
    const uint32_t hsize = 1u << P.hash_bits;
    const uint32_t hmask = hsize - 1u;

    float* qs = reinterpret_cast<float*>(smem);
    uint64_t* bk = reinterpret_cast<uint64_t*>(qs + dimpad);  // beam keys (bkey), 8-byte aligned
    uint32_t* hash = reinterpret_cast<uint32_t*>(bk + cap);
    float* sd = reinterpret_cast<float*>(hash + hsize);
    uint32_t* si = reinterpret_cast<uint32_t*>(sd + 64);
    // STOP / TRACE: the features (16 floats) and the network's hidden activations (64 floats)
    // use the stage, which holds nothing live between expansions
    float* feat = sd;
    float* hbuf = sd + 16;

    __shared__ uint32_t s_slot;
    // tables of 2^14 slots or more are sized to forget nothing (the distinct-evaluation count):
    // they probe; smaller ones take one exchange per neighbour
    const bool probe = TRACE || P.hash_bits >= 14;
    auto visit = [&](uint32_t* h, uint32_t m, uint32_t id) {
        return probe ? hash_insert<true>(h, m, id) : hash_insert<false>(h, m, id);
    };

    const uint32_t nq_all = P.nq_dev ? *P.nq_dev : P.nq;
    for (;;) {
        if (lane == 0) s_slot = atomicAdd(P.work, 1u);
        __syncthreads();
        const uint32_t slot = s_slot;
        __syncthreads();
        if (slot >= nq_all) break;
        const uint32_t qi = P.order ? P.order[slot] : slot;
        // per-query values are clamped to what the launch was sized for: k to the launch's k
        // (at most the row stride), max(ef, k) to the beam capacity
        int kq = P.k_pq ? P.k_pq[qi] : (int)P.k;
        kq = kq < 1 ? 1 : (kq > (int)P.k ? (int)P.k : kq);
        const int efq = P.ef_pq ? P.ef_pq[qi] : (int)P.ef;
        // the range search bounds its result set by ef alone (max_num_candidates)
        const int ef_want = (FILT && P.eps_mode) ? efq : (efq > kq ? efq : kq);
        // FILT keeps entries by the beam capacity (an overflow goes to a larger pass), so its
        // result-set bound stays unclamped
        const int ef_eff = (FILT || ef_want < (int)cap) ? ef_want : (int)cap;

        // query to LDS, hash cleared
        const float* qg = P.queries + (size_t)qi * (P.q_stride ? P.q_stride : g.dim);
        for (uint32_t i = lane; i < dimpad; i += 64) qs[i] = i < g.dim ? qg[i] : 0.f;
        for (uint32_t i = lane; i < hsize; i += 64) hash[i] = SENT;
        __syncthreads();

        int Hu = 0, Eu = 0, H0 = 0, E0 = 0;
        // ---- entry point and upper-layer greedy descent ----
        uint32_t cur = g.entry;
        if (lane == 0) si[0] = cur;
        __syncthreads();
        dist_batch<SPACE, MODE, NCH, U8>(g, qs, si, 1, sd, lane);
        __syncthreads();
        float curdist = sd[0];
        Eu = 1;
        for (int level = g.maxlevel; level > 0; level--) {
            bool changed = true;
            while (changed) {
                changed = false;
                const uint32_t* row = g.upper + ((size_t)g.upper_off[cur] + (size_t)(level - 1)) * g.maxM;
                Hu++;
                int cnt = 0;
                for (uint32_t b = 0; b < g.maxM; b += 64) {
                    uint32_t j = b + lane;
                    uint32_t id = j < g.maxM ? row[j] : SENT;
                    if (P.id_limit && id >= P.id_limit) id = SENT;
                    uint64_t m = ballot(id != SENT);
                    if (id != SENT) si[cnt + lanes_below(m, lane)] = id;
                    cnt += __popcll(m);
                }
                __syncthreads();
                if (cnt == 0) continue;
                dist_batch<SPACE, MODE, NCH, U8>(g, qs, si, cnt, sd, lane);
                __syncthreads();
                Eu += cnt;
                // first minimum in list order
                float bestd = INFINITY;
                int bestj = 0x7FFFFFFF;
                for (int j = lane; j < cnt; j += 64) {
                    float d = sd[j];
                    if (d < bestd) {
                        bestd = d;
                        bestj = j;
                    }
                }
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    float od = __shfl_xor(bestd, o);
                    int oj = __shfl_xor(bestj, o);
                    if (od < bestd || (od == bestd && oj < bestj)) {
                        bestd = od;
                        bestj = oj;
                    }
                }
                if (bestd < curdist) {
                    curdist = bestd;
                    cur = si[bestj];
                    changed = true;
                }
                __syncthreads();
            }
        }

        // ---- level 0 ----
        int n = 1;
        int nallowed = 1;  // allowed entries in the beam (FILT)
        if (FILT) {
            const uint32_t wdel = P.del_bits ? P.del_bits[cur >> 5] : 0u;
            const uint32_t wal = P.allow_bits ? P.allow_bits[cur >> 5] : ~0u;
            nallowed = id_allowed(wdel, wal, cur) ? 1 : 0;
        }
        if (lane == 0) {
            bk[0] = bkey(curdist, cur) | ((FILT && nallowed == 0) ? BD : 0ull);
            visit(hash, hmask, cur);
        }
        E0 = 1;
        __syncthreads();
        // segments of 2^sh entries, at most 64 of them over the most entries the beam can hold
        const int beam_max = FILT ? (int)cap : ef_eff;
        int sh = 0;
        while ((64 << sh) < beam_max) sh++;
        uint64_t head = beam_head(bk, n, sh, lane);
        uint64_t worst = bk[n - 1] & ~BFLAGS;  // the last beam entry's key (the admission bound)
        int scan_lo = 0;
        StopState st;
        st.d_entry = curdist;
        st.improved_recent = 0;
        st.last_change = 0;
        st.win_pos = 1 << 30;
        st.prev_pos[0] = st.prev_pos[1] = st.prev_pos[2] = 1 << 30;
        if (STOP || TRACE) {
            float s = 0.f;
            for (uint32_t i = lane; i < g.dim; i += 64) s = fma_(qs[i], qs[i], s);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
            st.qnorm = sqrtf(s);
        }
        int nrec = 0;
        bool stopped = false;
        // the 12 traversal features of the current state into feat[] (STOP / TRACE)
        auto make_features = [&]() { beam_features(bk, n, scan_lo, kq, ef_eff, H0, st, feat, lane); };
        const int hcap = (P.hop_cap && P.hop_cap[qi] > 0) ? P.hop_cap[qi] : 0x7FFFFFFF;
        bool ovf = false;  // FILT: the beam overflowed (the result may differ from the reference's)
        // Link-row prefetch (maxM0 <= 32): lanes 32..63 load the row of the node expected to be
        // expanded next, pf_node; the next expansion takes it from them when it is that node.
        uint32_t pf_node = SENT, pf_row = SENT;
        auto prefetch_row = [&](uint32_t nd) {
            uint32_t v = SENT;
            if (lane >= 32 && nd != SENT && (lane & 31) < (int)g.maxM0) v = g.links0[(size_t)nd * g.maxM0 + (lane & 31)];
            pf_row = v;
            pf_node = nd;
        };
        auto plan_next = [&](uint32_t next) { prefetch_row(next); };
        bool lazy = false;  // the plain and learned kernels run their own level-0 loop (dim % 16 == 0)
        if constexpr (MODE == 1 && NCH > 0 && !TRACE && !FILT) {
            if (g.maxM0 <= 32) {
                level0_lazy<SPACE, NCH, STOP, U8>(g, qs, bk, hash, hmask, probe, reinterpret_cast<uint64_t*>(sd), ef_eff,
                                                  hcap, lane, n, H0, E0, scan_lo, P, st, kq, feat, hbuf, stopped);
                lazy = true;
            }
        }
        if (!lazy) for (;;) {
            // first unexpanded entry at or after scan_lo, and the one after it (node2, key k2):
            // the next expansion unless a new key of this expansion lands before it
            int p = -1;
            uint32_t node2 = SENT;
            uint64_t k2 = ~0ull, bkp = 0;
            for (int b = scan_lo; b < n; b += 64) {
                int i = b + lane;
                const int ic = i < n ? i : n - 1;  // every lane reads; the ballot drops i >= n
                const uint64_t e = bk[ic];
                const bool un = (i < n) & ((e & BX) == 0);
                uint64_t m = ballot(un);
                if (m) {
                    const int l1 = __ffsll((unsigned long long)m) - 1;
                    p = b + l1;
                    bkp = readlane64(e, l1);
                    const uint64_t m2 = m & (m - 1);
                    if (m2) {
                        k2 = readlane64(e, __ffsll((unsigned long long)m2) - 1) & ~BFLAGS;
                        node2 = bid(k2);
                    }
                    break;
                }
            }
            if (p < 0 || H0 >= hcap) break;
            if (FILT && P.eps_mode) {
                // EpsilonSearchStopCondition::should_stop_search (stop_condition.h:232-243); a
                // disallowed entry point is the reference's candidate at FLT_MAX
                const float cd = (H0 == 0 && (bkp & BD)) ? __FLT_MAX__ : bdist(bkp);
                if (cd > P.eps && nallowed >= P.eps_min) break;
            }
            const uint32_t node = bid(bkp);
            __syncthreads();
            if (lane == 0) bk[p] = bkp | BX;
            H0++;

            // neighbours of node: keep those not visited, in list order
            int cnt = 0;
            if (g.maxM0 <= 32) {
                // lanes 0..31 take node's row: from the prefetch when it was this node's, else
                // loaded now
                const int lr = lane & 31;
                const bool upper = lane >= 32;
                uint32_t cur = (uint32_t)__shfl_xor((int)pf_row, 32);
                if (node != pf_node) cur = (!upper && lr < (int)g.maxM0) ? g.links0[(size_t)node * g.maxM0 + lr] : SENT;
                const uint32_t id = (!upper && lr < (int)g.maxM0) ? cur : SENT;
                const bool fresh = id != SENT && visit(hash, hmask, id);
                const uint64_t m = ballot(fresh);
                if (fresh) si[lanes_below(m, lane)] = id;
                cnt = __popcll(m);
            } else {
                const uint32_t* row = g.links0 + (size_t)node * g.maxM0;
                for (uint32_t b = 0; b < g.maxM0; b += 64) {
                    uint32_t j = b + lane;
                    uint32_t id = j < g.maxM0 ? row[j] : SENT;
                    bool fresh = id != SENT && visit(hash, hmask, id);
                    uint64_t m = ballot(fresh);
                    if (fresh) si[cnt + lanes_below(m, lane)] = id;
                    cnt += __popcll(m);
                }
            }
            __syncthreads();
            // the batch: distances, admission, insertion points and ranks; m > 0 keys to merge
            int m = 0, r = 0, pos = 0, spos = 0, first_pos = 0;
            bool ok = false;
            uint64_t key = ~0ull;   // the staged key (bkey), ~0 when not admitted
            uint32_t idf = SENT;    // id with its DFLAG (FILT)
            uint64_t kmin = ~0ull;  // the smallest key to merge
            if (cnt > 0) {
                // FILT: the membership words of the fresh ids load with the vectors
                uint32_t fw_del = 0u, fw_al = ~0u;
                if (FILT && lane < cnt) {
                    const uint32_t fid = si[lane];
                    if (P.del_bits) fw_del = P.del_bits[fid >> 5];
                    if (P.allow_bits) fw_al = P.allow_bits[fid >> 5];
                }
                dist_batch<SPACE, MODE, NCH, U8>(g, qs, si, cnt, sd, lane);
                __syncthreads();
                if (FILT) {
                    if (lane < cnt) {
                        const uint32_t fid = si[lane];
                        if (!id_allowed(fw_del, fw_al, fid)) si[lane] = fid | DFLAG;
                    }
                    __syncthreads();
                }
                E0 += cnt;
                // admission against the current worst (when full)
                if (lane < cnt) {
                    idf = si[lane];
                    key = bkey(sd[lane], idf & ID_MASK);
                    const bool full = FILT ? nallowed >= ef_eff : n >= ef_eff;
                    if (full && !(key < worst)) key = ~0ull;
                }
                const uint64_t live = ballot(key != ~0ull);
                if (live) {
                    // Drop nodes already in the beam (re-evaluated after a hash overwrite).  pos =
                    // number of beam keys below this key: its insertion point, located through the
                    // segment heads when few keys are staged (one ballot per key), else by a
                    // full-depth search per lane.
                    ok = key != ~0ull;
                    pos = __popcll(live) <= kHeadKeys ? beam_lower_bound_heads(bk, n, sh, head, key, live, lane)
                                                      : (ok ? beam_lower_bound64(bk, n, key) : 0);
                    if (ok && pos < n && (bk[pos] & ~BFLAGS) == key) ok = false;
                    const uint64_t km = ballot(ok);
                    m = __popcll(km);
                    if (m > 0) {
                        // Rank of each staged key among the staged keys (keys are distinct):
                        // a scalar loop over the staged lanes, no cross-lane sort network.
                        const uint32_t klo = (uint32_t)key, khi = (uint32_t)(key >> 32);
                        for (uint64_t t = km; t; t &= t - 1) {
                            const int L = __ffsll((unsigned long long)t) - 1;
                            const uint64_t kl = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)khi, L) << 32) |
                                                (uint32_t)__builtin_amdgcn_readlane((int)klo, L);
                            r += kl < key ? 1 : 0;
                        }
                        // insertion points in rank order: lane r holds the point of rank r (the
                        // other lanes park their value in lanes m..63)
                        const int dst = ok ? r : m + lanes_below(~km, lane);
                        spos = __builtin_amdgcn_ds_permute(dst * 4, pos);
                        first_pos = __builtin_amdgcn_readlane(spos, 0);
                        const int L0 = __ffsll((unsigned long long)ballot(ok && r == 0)) - 1;
                        kmin = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)khi, L0) << 32) |
                               (uint32_t)__builtin_amdgcn_readlane((int)klo, L0);
                    }
                }
            }
            // the next expansion is node2 or the smallest new key, whichever is smaller: its row
            // loads while the beam moves
            if (g.maxM0 <= 32) plan_next(kmin < k2 ? bid(kmin) : node2);
            if (m > 0) {
                // Move beam entries [first_pos, n) right, highest chunks first, eight chunks per
                // batch with every read of a batch issued before its writes.  Entry i moves by the
                // number of staged insertion points <= i.  A batch writes only at or above its
                // lowest chunk base, which the next (lower) batch never reads, and a wave's LDS
                // operations execute in order.
                const int lim = FILT ? (int)cap : ef_eff;  // entries kept by position
                const int new_n = n + m < lim ? n + m : lim;
                ovf |= FILT && n + m > lim;
                const int hi_chunk = (n - 1) & ~63;
                const int lo_chunk = first_pos & ~63;
                // Reads are unconditional (indices below cap, clamped at 0); a lane whose entry
                // does not move writes to its own 8-byte slot of the distance-staging area (sd and
                // si, 512 contiguous bytes holding no live data during the merge), so no lane
                // branches.
                const int trash = (int)(reinterpret_cast<uint64_t*>(sd) - bk) + lane;
                // the rank lanes of spos (the insertion points, non-decreasing in rank)
                const uint64_t rank_lanes = m >= 64 ? ~0ull : ((1ull << m) - 1ull);
                for (int b0 = hi_chunk; b0 >= lo_chunk; b0 -= 8 * 64) {
                    uint64_t vk[8];
                    int mv[8];
                    const int base = b0 + lane;
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int i = base - 64 * u;
                        vk[u] = bk[i > 0 ? i : 0];
                    }
                    // shift of entry i = insertion points <= i.  Per chunk [cb, cb + 64): lo
                    // points lie at or below cb and move the whole chunk; only the points inside
                    // the chunk (ranks lo..hi-1) are compared per lane
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int cb = b0 - 64 * u;
                        const int lo = __popcll(__builtin_amdgcn_ballot_w64(spos <= cb) & rank_lanes);
                        const int hi = __popcll(__builtin_amdgcn_ballot_w64(spos <= cb + 63) & rank_lanes);
                        int sv = lo;
                        for (int rr = lo; rr < hi; rr++)
                            sv += __builtin_amdgcn_readlane(spos, rr) <= base - 64 * u ? 1 : 0;
                        mv[u] = sv;
                    }
                    __asm__ volatile("" ::: "memory");
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int i = base - 64 * u;
                        const int np = i + mv[u];
                        const bool w = (b0 - 64 * u >= lo_chunk) & (i >= first_pos) & (i < n) & (np < lim);
                        bk[w ? np : trash] = vk[u];
                    }
                    __asm__ volatile("" ::: "memory");
                }
                const int fp = pos + r;
                if (ok && fp < lim) bk[fp] = key | ((FILT && (idf & DFLAG)) ? BD : 0ull);
                const int min_fp = first_pos;
                // Every position below min(first insert, p) holds an old entry whose old index
                // was below p, hence already expanded.
                scan_lo = min_fp < p ? min_fp : p;
                if (STOP || TRACE) {
                    int kk = kq < new_n ? kq : new_n;
                    if (min_fp < kk) {
                        st.improved_recent |= 1;
                        st.last_change = H0;
                    }
                    st.win_pos = min_fp < st.win_pos ? min_fp : st.win_pos;
                }
                n = new_n;
                __syncthreads();
                if (FILT) {
                    nallowed += __popcll(ballot(ok && (idf & DFLAG) == 0u));
                    if (nallowed >= ef_eff) {
                        // end the beam right after the ef-th allowed entry: drop the last
                        // (nallowed - ef) allowed entries and whatever lies beyond
                        int remaining = nallowed - ef_eff, newn = n;
                        for (int b = (n - 1) & ~63; b >= 0; b -= 64) {
                            const int i = b + lane;
                            const uint64_t am = ballot(i < n && (bk[i] & BD) == 0u);
                            const int c = __popcll(am);
                            if (c > remaining) {
                                uint64_t t = am;
                                for (int j = 1; j < c - remaining; j++) t &= t - 1;
                                newn = b + __ffsll((unsigned long long)t);
                                break;
                            }
                            remaining -= c;
                        }
                        n = newn;
                        nallowed = ef_eff;
                        scan_lo = scan_lo < n ? scan_lo : n;
                    }
                }
                head = beam_head(bk, n, sh, lane);
                worst = bk[n - 1] & ~BFLAGS;
            } else {
                scan_lo = p;
            }
            if (STOP || TRACE) {
                const int period = STOP ? P.stop.period : P.trace_period;
                if (H0 % period == 0) {
                    make_features();
                    st.prev_pos[2] = st.prev_pos[1];
                    st.prev_pos[1] = st.prev_pos[0];
                    st.prev_pos[0] = st.win_pos < ef_eff ? st.win_pos : ef_eff;
                    st.win_pos = 1 << 30;
                    __syncthreads();
                    if (TRACE && nrec < P.trace_max_rec) {
                        float* fo = P.trace_feat + ((size_t)qi * P.trace_max_rec + nrec) * kStopFeatures;
                        if (lane < kStopFeatures) fo[lane] = feat[lane];
                        uint32_t* to = P.trace_topk + ((size_t)qi * P.trace_max_rec + nrec) * P.k_stride;
                        for (int i = lane; i < kq; i += 64) to[i] = i < n ? bid(bk[i]) : SENT;
                        nrec++;
                    }
                    if (STOP && H0 >= P.stop.min_hops) {
                        if (lane < kStopFeatures)
                            feat[lane] = (feat[lane] - P.stop.mean[lane]) * P.stop.inv_std[lane];
                        __syncthreads();
                        float logit = mlp_logit(P.stop, feat, hbuf, lane);
                        if (logit >= P.stop.logit_threshold) stopped = true;
                    }
                    __syncthreads();
                }
                st.improved_recent <<= 1;
                if (stopped) break;
            }
            // a pass that lists overflowing queries for a larger beam abandons them at once
            if (FILT && ovf && P.ovf_list) break;
        }

        // where the search ended by its own rule (no stop, no hop cap): a trace records the
        // state; the first pass of an extended learned search asks the network whether the
        // top-k has converged and, when it is not confident, lists the query for the larger ef
        if ((STOP && P.ext_list) || (TRACE && P.trace_end)) {
            if (!stopped && H0 < hcap) {
                make_features();
                __syncthreads();
                if (TRACE && nrec < P.trace_max_rec) {
                    float* fo = P.trace_feat + ((size_t)qi * P.trace_max_rec + nrec) * kStopFeatures;
                    if (lane < kStopFeatures) fo[lane] = feat[lane];
                    uint32_t* to = P.trace_topk + ((size_t)qi * P.trace_max_rec + nrec) * P.k_stride;
                    for (int i = lane; i < kq; i += 64) to[i] = i < n ? bid(bk[i]) : SENT;
                    nrec++;
                }
                if (STOP) {
                    if (lane < kStopFeatures)
                        feat[lane] = (feat[lane] - P.stop.mean[lane]) * P.stop.inv_std[lane];
                    __syncthreads();
                    const float logit = mlp_logit(P.stop, feat, hbuf, lane);
                    if (logit < P.stop.ext_logit_threshold) {
                        if (lane == 0) P.ext_list[atomicAdd(P.ext_count, 1u)] = qi;
                        __syncthreads();
                        continue;
                    }
                }
                __syncthreads();
            }
        }
        if (FILT && ovf && P.ovf_list) {
            if (lane == 0) P.ovf_list[atomicAdd(P.ovf_count, 1u)] = qi;
            __syncthreads();
            continue;
        }
        // entries kept for the output: k, or for the range search the whole result set
        const int keep = (FILT && P.eps_mode) ? ef_eff : kq;
        if (FILT) {
            // the result set: allowed entries in beam order, compacted to the front in place
            // (a chunk's reads precede its writes, and every entry moves down or stays)
            int got = 0;
            for (int b = 0; b < n && got < keep; b += 64) {
                const int i = b + lane;
                const uint64_t kv = i < n ? bk[i] : BD;
                const bool al = i < n && (kv & BD) == 0u;
                const uint64_t am = ballot(al);
                const int slot = got + lanes_below(am, lane);
                __asm__ volatile("" ::: "memory");
                if (al && slot < keep) bk[slot] = kv;
                __asm__ volatile("" ::: "memory");
                got += __popcll(am);
            }
            n = got < keep ? got : keep;
            if (lane == 0 && ovf) atomicOr(P.status, 1u);
            __syncthreads();
        }
        // ---- output: rows ordered by (distance, label) ----
        // The beam is sorted by (distance, id): entry i's rank by (distance, label) is the start
        // of its run of equal distances plus the run entries with a smaller (label, position).
        int nout = kq < n ? kq : n;
        int row_len = kq;
        if (FILT && P.eps_mode) {
            // filter_results (stop_condition.h:262-269): the results within eps, a prefix
            int neps = 0;
            for (int b = 0; b < n; b += 64) {
                const int i = b + lane;
                const uint64_t m = ballot(i < n && bdist(bk[i]) <= P.eps);
                neps += __popcll(m);
                if (m != ~0ull) break;
            }
            nout = neps;
            row_len = (int)P.k_stride;
            if (lane == 0) P.counts[qi] = (uint32_t)neps;
        }
        __syncthreads();
        uint64_t* ol = P.out_labels + (size_t)qi * P.k_stride;
        float* od = P.out_dists + (size_t)qi * P.k_stride;
        for (int i = lane; i < nout; i += 64) {
            const uint64_t ki = bk[i];
            const float di = bdist(ki);
            const uint32_t oi = (uint32_t)(ki >> 32);  // the distance's order bits: equal iff equal
            const uint64_t li = g.labels[bid(ki)];
            int s0 = i;
            while (s0 > 0 && (uint32_t)(bk[s0 - 1] >> 32) == oi) s0--;
            int r = s0;
            for (int j = s0; j < nout && (uint32_t)(bk[j] >> 32) == oi; j++) {
                if (j == i) continue;
                const uint64_t lj = g.labels[bid(bk[j])];
                r += (lj < li || (lj == li && j < i)) ? 1 : 0;
            }
            if (r < row_len) {
                ol[r] = li;
                od[r] = di;
            }
        }
        for (int i = (nout < row_len ? nout : row_len) + lane; i < row_len; i += 64) {
            ol[i] = ~0ull;
            od[i] = INFINITY;
        }
        if (P.counters && lane == 0) {
            int32_t* c = P.counters + (size_t)qi * 4;
            c[0] = H0;
            c[1] = E0;
            c[2] = Hu;
            c[3] = Eu;
        }
        if (TRACE) {
            if (lane == 0) P.trace_nrec[qi] = nrec;
            uint32_t* tf = P.trace_final + (size_t)qi * P.k_stride;
            for (int i = lane; i < kq; i += 64) tf[i] = i < n ? bid(bk[i]) : SENT;
        }
        __syncthreads();
    }
}

// Plain streaming copy (16 bytes per lane per step): the reference kernel of known traffic
// used to calibrate the FETCH_SIZE / WRITE_SIZE hardware counters.
__global__ __launch_bounds__(256) void copy_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n16) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) dst[i] = src[i];
}

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes % 16 != 0) return hipErrorInvalidValue;
    size_t n16 = bytes / 16;
    size_t blocks = (n16 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(copy_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, (uint4*)dst, (const uint4*)src, n16);
    return hipGetLastError();
}

// labels[i] = external label of internal id ids[i] (SENT -> all ones).
__global__ void id_labels_kernel(const uint64_t* __restrict__ lab, uint32_t n_nodes, const uint32_t* __restrict__ ids,
                                 size_t n, uint64_t* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = ids[i] & ID_MASK;
    out[i] = (ids[i] != SENT && id < n_nodes) ? lab[id] : ~0ull;
}

hipError_t launch_id_labels(const DevGraph& g, const uint32_t* ids, size_t n, uint64_t* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(id_labels_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g.labels, g.n, ids, n, out);
    return hipGetLastError();
}

// One output word per thread and step (grid-stride): the four coordinates 16t+4g..16t+4g+3 of row
// r as bytes, placed at word g * (dim/16) + t of the coded row.  A coordinate whose FP32 bits are
// not those of an integer in 0..255 (a fraction, out of range, -0, NaN) sets the flag, once per
// wave that met one.
__global__ __launch_bounds__(256) void code_u8_kernel(const float* __restrict__ vec, uint32_t n, uint32_t dim,
                                                      uint32_t stride, uint32_t* __restrict__ out,
                                                      uint32_t* __restrict__ flag) {
    const uint32_t words = dim / 4, nch = dim / 16;
    const size_t total = (size_t)n * words;
    bool ok = true;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t r = (uint32_t)(i / words), wi = (uint32_t)(i % words);
        const uint32_t g = wi / nch, t = wi % nch;
        const float4 x = *reinterpret_cast<const float4*>(vec + (size_t)r * stride + 16 * t + 4 * g);
        const float xs[4] = {x.x, x.y, x.z, x.w};
        uint32_t w = 0u;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float f = xs[j];
            const bool in = f >= 0.f && f <= 255.f;  // false for NaN
            const uint32_t b = in ? (uint32_t)f : 0u;
            ok &= in && __float_as_uint((float)b) == __float_as_uint(f);
            w |= b << (8 * j);
        }
        out[i] = w;
    }
    const uint64_t bad = __ballot(!ok);
    if (bad && (threadIdx.x & 63) == (unsigned)(__ffsll((unsigned long long)bad) - 1)) atomicOr(flag, 1u);
}

hipError_t launch_code_u8(const float* vec, uint32_t n, uint32_t dim, uint32_t stride, uint8_t* out, uint32_t* flag,
                          hipStream_t s) {
    if (!search_fixed_chunks(dim) || stride % 4 != 0) return hipErrorInvalidValue;
    const size_t words = (size_t)n * (dim / 4);
    if (words == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((words + 255) / 256, 4096);
    hipLaunchKernelGGL(code_u8_kernel, dim3((unsigned)blocks), dim3(256), 0, s, vec, n, dim, stride,
                       reinterpret_cast<uint32_t*>(out), flag);
    return hipGetLastError();
}

__global__ void zero_u32_kernel(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

hipError_t launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 63) / 64), dim3(64), 0, s, p, n);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Host-side launch
// ------------------------------------------------------------------------------------------
size_t search_lds_bytes(uint32_t dim, uint32_t cap, uint32_t hash_bits) {
    uint32_t dimpad = (dim + 3u) & ~3u;
    return (size_t)dimpad * 4 + (size_t)cap * 8 + ((size_t)1 << hash_bits) * 4 + 64 * 8;
}

template <int SPACE, int MODE, int NCH, bool STOP, bool TRACE, bool FILT, bool U8 = false>
static hipError_t launch_t(const SearchParams& p, int blocks, size_t lds, hipStream_t s) {
    auto k = hnsw_search_kernel<SPACE, MODE, NCH, STOP, TRACE, FILT, U8>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, dim3(blocks), dim3(64), lds, s, p);
    return hipGetLastError();
}

// Chunk counts compiled with a fixed unroll (dim = 16, 32, 64, 96, 128); the 8-bit coded rows
// are read only by these.
bool search_fixed_chunks(uint32_t dim) {
    if (dim % 16 != 0) return false;
    const uint32_t c = dim / 16;
    return c == 1 || c == 2 || c == 4 || c == 6 || c == 8;
}

template <int SPACE, int NCH, bool STOP, bool TRACE, bool FILT>
static hipError_t launch_code(const SearchParams& p, int blocks, size_t lds, hipStream_t s) {
    return p.g.vec8 ? launch_t<SPACE, 1, NCH, STOP, TRACE, FILT, true>(p, blocks, lds, s)
                    : launch_t<SPACE, 1, NCH, STOP, TRACE, FILT, false>(p, blocks, lds, s);
}

template <int SPACE, bool STOP, bool TRACE, bool FILT = false>
static hipError_t launch_dim(const SearchParams& p, int blocks, size_t lds, hipStream_t s) {
    if (p.g.vec8 && !search_fixed_chunks(p.g.dim)) return hipErrorInvalidValue;
    if (p.g.dim % 16 != 0) return launch_t<SPACE, 0, 0, STOP, TRACE, FILT>(p, blocks, lds, s);
    switch (p.g.dim / 16) {
        case 1: return launch_code<SPACE, 1, STOP, TRACE, FILT>(p, blocks, lds, s);
        case 2: return launch_code<SPACE, 2, STOP, TRACE, FILT>(p, blocks, lds, s);
        case 4: return launch_code<SPACE, 4, STOP, TRACE, FILT>(p, blocks, lds, s);
        case 6: return launch_code<SPACE, 6, STOP, TRACE, FILT>(p, blocks, lds, s);
        case 8: return launch_code<SPACE, 8, STOP, TRACE, FILT>(p, blocks, lds, s);
        default: return launch_t<SPACE, 1, 0, STOP, TRACE, FILT>(p, blocks, lds, s);
    }
}

hipError_t launch_search(const SearchParams& p, int blocks, hipStream_t s) {
    bool stop = p.use_stop != 0, trace = p.trace_feat != nullptr;
    const bool filt = p.del_bits || p.allow_bits || p.eps_mode;
    size_t lds = search_lds_bytes(p.g.dim, p.cap, p.hash_bits);
    uint32_t* w = p.work;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, w, 1u);
    if (filt)
        return p.g.space == 0 ? launch_dim<0, false, false, true>(p, blocks, lds, s)
                              : launch_dim<1, false, false, true>(p, blocks, lds, s);
    if (p.g.space == 0) {
        if (trace) return launch_dim<0, false, true>(p, blocks, lds, s);
        if (stop) return launch_dim<0, true, false>(p, blocks, lds, s);
        return launch_dim<0, false, false>(p, blocks, lds, s);
    }
    if (trace) return launch_dim<1, false, true>(p, blocks, lds, s);
    if (stop) return launch_dim<1, true, false>(p, blocks, lds, s);
    return launch_dim<1, false, false>(p, blocks, lds, s);
}

int search_blocks_per_cu(const SearchParams& p, bool filtered) {
    size_t lds = search_lds_bytes(p.g.dim, p.cap, p.hash_bits);
    int nb = 0;
    hipError_t e;
    const bool d128 = p.g.dim == 128 && p.g.space == 0, u8 = p.g.vec8 != nullptr;
    if (filtered && d128)
        e = u8 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, hnsw_search_kernel<0, 1, 8, false, false, true, true>, 64, lds)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, hnsw_search_kernel<0, 1, 8, false, false, true, false>, 64, lds);
    else if (filtered)
        e = p.g.space == 0
                ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, hnsw_search_kernel<0, 1, 0, false, false, true, false>, 64, lds)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, hnsw_search_kernel<1, 1, 0, false, false, true, false>, 64, lds);
    else if (d128)
        e = u8 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, hnsw_search_kernel<0, 1, 8, false, false, false, true>, 64, lds)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, hnsw_search_kernel<0, 1, 8, false, false, false, false>, 64, lds);
    else if (p.g.space == 0)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, hnsw_search_kernel<0, 1, 0, false, false, false, false>, 64, lds);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, hnsw_search_kernel<1, 1, 0, false, false, false, false>, 64, lds);
    if (e != hipSuccess || nb <= 0) nb = 1;
    return nb;
}

}  // namespace hg
