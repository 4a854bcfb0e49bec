// Multi-vector search on the GPU: searchStopConditionClosest with
// MultiVectorSearchStopCondition (hnswalg.h:1326-1377, stop_condition.h:82-200) in (distance,
// id) order.  One wave per query, persistent workgroups as in the query kernel.
//
// The result set is bounded by its number of distinct documents (ef_coll): a candidate is
// admitted while fewer documents are held or when it is closer than the worst result, and after
// each admission the farthest results leave the set while it holds more than ef_coll documents.
// Both steps depend on the order of arrivals, so the neighbours of an expansion are taken one at
// a time in list order (not set-wise as in the query kernel).  The beam (sorted by (distance,
// id), entries flagged expanded / not in the result set) holds every admitted candidate; once
// ef_coll documents are held, nothing beyond the worst result can be expanded again and the
// beam ends there.  Document counts live in a small LDS table (one entry per held document).
// At the end the farthest results leave until num_docs documents remain (filter_results).
#include "common.h"
#include "dist.h"
#include "kernels.h"
#include "wave.h"

namespace hg {

namespace {

__device__ __forceinline__ bool mv_allowed(const MvParams& P, uint32_t id) {
    const uint32_t wdel = P.del_bits ? P.del_bits[id >> 5] : 0u;
    const uint32_t wal = P.allow_bits ? P.allow_bits[id >> 5] : ~0u;
    const uint32_t b = id & 31u;
    return ((wdel >> b) & 1u) == 0u && ((wal >> b) & 1u) != 0u;
}

// Index of the last beam entry in the result set (not flagged), or -1.
__device__ __forceinline__ int mv_last_result(const uint32_t* bi, int n, int lane) {
    for (int b = (n - 1) & ~63; b >= 0; b -= 64) {
        const int i = b + lane;
        const uint64_t m = ballot(i < n && (bi[i] & DFLAG) == 0u);
        if (m) return b + 63 - __clzll((long long)m);
    }
    return -1;
}

// Slot of doc in the table [0, nd), or -1.
__device__ __forceinline__ int mv_find(const uint32_t* dkey, int nd, uint32_t doc, int lane) {
    for (int b = 0; b < nd; b += 64) {
        const int i = b + lane;
        const uint64_t m = ballot(i < nd && dkey[i] == doc);
        if (m) return b + __ffsll((unsigned long long)m) - 1;
    }
    return -1;
}

}  // namespace

template <int SPACE, int MODE, int NCH>
__global__ __launch_bounds__(64) void hnsw_mv_kernel(MvParams P) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const DevGraph& g = P.g;
    const uint32_t dimpad = (g.dim + 3u) & ~3u;
    const int cap = (int)P.cap;
    const uint32_t hsize = 1u << P.hash_bits, hmask = hsize - 1u;
    float* qs = reinterpret_cast<float*>(smem);
    float* bd = qs + dimpad;
    uint32_t* bi = reinterpret_cast<uint32_t*>(bd + cap);
    uint32_t* hash = bi + cap;
    float* sd = reinterpret_cast<float*>(hash + hsize);
    uint32_t* si = reinterpret_cast<uint32_t*>(sd + 64);
    uint32_t* dkey = si + 64;
    uint32_t* dcnt = dkey + P.dcap;
    __shared__ uint32_t s_slot;
    const int efc = (int)P.ef_coll;

    for (;;) {
        if (lane == 0) s_slot = atomicAdd(P.work, 1u);
        __syncthreads();
        const uint32_t qi = s_slot;
        __syncthreads();
        if (qi >= P.nq) break;
        const float* qg = P.queries + (size_t)qi * g.dim;
        for (uint32_t i = lane; i < dimpad; i += 64) qs[i] = i < g.dim ? qg[i] : 0.f;
        for (uint32_t i = lane; i < hsize; i += 64) hash[i] = SENT;
        __syncthreads();

        int Hu = 0, Eu = 0, H0 = 0, E0 = 0;
        // entry point and upper-level greedy descent (hnswalg.h:1335-1360)
        uint32_t cur = g.entry;
        if (lane == 0) si[0] = cur;
        __syncthreads();
        dist_batch<SPACE, MODE, NCH>(g, qs, si, 1, sd, lane);
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
                    const uint32_t j = b + lane;
                    const uint32_t id = j < g.maxM ? row[j] : SENT;
                    const uint64_t m = ballot(id != SENT);
                    if (id != SENT) si[cnt + lanes_below(m, lane)] = id;
                    cnt += __popcll(m);
                }
                __syncthreads();
                if (cnt == 0) continue;
                dist_batch<SPACE, MODE, NCH>(g, qs, si, cnt, sd, lane);
                __syncthreads();
                Eu += cnt;
                float bestd = INFINITY;
                int bestj = 0x7FFFFFFF;
                for (int j = lane; j < cnt; j += 64) {
                    const float d = sd[j];
                    if (d < bestd) {
                        bestd = d;
                        bestj = j;
                    }
                }
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    const float od = __shfl_xor(bestd, o);
                    const int oj = __shfl_xor(bestj, o);
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

        // level 0
        int n = 1, nd = 0;
        const bool ep_ok = mv_allowed(P, cur);
        if (lane == 0) {
            bd[0] = curdist;
            bi[0] = cur | (ep_ok ? 0u : DFLAG);
            hash_insert(hash, hmask, cur);
            if (ep_ok) {
                dkey[0] = P.docs[cur];
                dcnt[0] = 1u;
            }
        }
        if (ep_ok) nd = 1;
        E0 = 1;
        __syncthreads();
        int scan_lo = 0;
        bool ovf = false;
        for (;;) {
            int p = -1;
            for (int b = scan_lo; b < n; b += 64) {
                const int i = b + lane;
                const uint64_t m = ballot(i < n && (bi[i] & XFLAG) == 0u);
                if (m) {
                    p = b + __ffsll((unsigned long long)m) - 1;
                    break;
                }
            }
            if (p < 0) break;
            const uint32_t bip = bi[p];
            const uint32_t node = bip & ID_MASK;
            __syncthreads();
            if (lane == 0) bi[p] = bip | XFLAG;
            H0++;
            // fresh neighbours in list order
            int cnt = 0;
            const uint32_t* row = g.links0 + (size_t)node * g.maxM0;
            for (uint32_t b = 0; b < g.maxM0; b += 64) {
                const uint32_t j = b + lane;
                const uint32_t id = j < g.maxM0 ? row[j] : SENT;
                const bool fresh = id != SENT && hash_insert(hash, hmask, id);
                const uint64_t m = ballot(fresh);
                if (fresh) si[cnt + lanes_below(m, lane)] = id;
                cnt += __popcll(m);
            }
            __syncthreads();
            int min_pos = p;
            if (cnt > 0) {
                bool al = false;
                if (lane < cnt) al = mv_allowed(P, si[lane]);
                dist_batch<SPACE, MODE, NCH>(g, qs, si, cnt, sd, lane);
                __syncthreads();
                if (lane < cnt && !al) si[lane] |= DFLAG;
                __syncthreads();
                E0 += cnt;
                for (int j = 0; j < cnt; j++) {
                    const float d = sd[j];
                    const uint32_t idf = si[j], id = idf & ID_MASK;
                    if (nd >= efc) {  // held documents at the bound: only keys below the worst result
                        const int lr = mv_last_result(bi, n, lane);
                        if (lr >= 0 && !key_less(d, id, bd[lr], bi[lr] & ID_MASK)) continue;
                    }
                    const int pos = beam_lower_bound(bd, bi, n, d, id);
                    if (pos < n && bd[pos] == d && (bi[pos] & ID_MASK) == id) continue;  // re-evaluated
                    if (pos >= cap) {
                        ovf = true;
                        continue;
                    }
                    // shift [pos, n) up by one (the last entry drops when the beam is full)
                    const int newn = n < cap ? n + 1 : cap;
                    if (n >= cap) ovf = true;
                    for (int c = (newn - 1) & ~63; c >= (pos & ~63); c -= 64) {
                        const int i = c + lane;
                        const bool mv = i > pos && i < newn;
                        float vd = 0.f;
                        uint32_t vi = 0u;
                        if (mv) {
                            vd = bd[i - 1];
                            vi = bi[i - 1];
                        }
                        __asm__ volatile("" ::: "memory");
                        if (mv) {
                            bd[i] = vd;
                            bi[i] = vi;
                        }
                        __asm__ volatile("" ::: "memory");
                    }
                    if (lane == 0) {
                        bd[pos] = d;
                        bi[pos] = idf;
                    }
                    n = newn;
                    min_pos = pos < min_pos ? pos : min_pos;
                    __syncthreads();
                    if ((idf & DFLAG) == 0u) {  // add_point_to_result
                        const uint32_t doc = P.docs[id];
                        const int k = mv_find(dkey, nd, doc, lane);
                        if (lane == 0) {
                            if (k >= 0) {
                                dcnt[k] += 1u;
                            } else {
                                dkey[nd] = doc;
                                dcnt[nd] = 1u;
                            }
                        }
                        if (k < 0) nd++;
                        __syncthreads();
                        // remove_point_from_result of the farthest results (stop_condition.h:118-125)
                        while (nd > efc) {
                            const int lr = mv_last_result(bi, n, lane);
                            if (lr < 0) break;
                            const uint32_t rid = bi[lr] & ID_MASK;
                            const uint32_t rdoc = P.docs[rid];
                            const int kk = mv_find(dkey, nd, rdoc, lane);
                            const bool gone = kk >= 0 && dcnt[kk] == 1u;
                            __syncthreads();
                            if (lane == 0) {
                                bi[lr] |= DFLAG;
                                if (kk >= 0) {
                                    if (gone) {
                                        dkey[kk] = dkey[nd - 1];
                                        dcnt[kk] = dcnt[nd - 1];
                                    } else {
                                        dcnt[kk] -= 1u;
                                    }
                                }
                            }
                            if (gone) nd--;
                            __syncthreads();
                        }
                    }
                }
                if (nd >= efc) {
                    // nothing beyond the worst result can be expanded or return
                    const int lr = mv_last_result(bi, n, lane);
                    n = lr + 1;
                }
            }
            scan_lo = min_pos < n ? min_pos : n;
        }
        // filter_results (stop_condition.h:169-185): the farthest results leave until num_docs
        // documents remain
        while (nd > (int)P.num_docs) {
            const int lr = mv_last_result(bi, n, lane);
            if (lr < 0) break;
            const uint32_t rdoc = P.docs[bi[lr] & ID_MASK];
            const int kk = mv_find(dkey, nd, rdoc, lane);
            const bool gone = kk >= 0 && dcnt[kk] == 1u;
            __syncthreads();
            if (lane == 0) {
                bi[lr] |= DFLAG;
                if (kk >= 0) {
                    if (gone) {
                        dkey[kk] = dkey[nd - 1];
                        dcnt[kk] = dcnt[nd - 1];
                    } else {
                        dcnt[kk] -= 1u;
                    }
                }
            }
            if (gone) nd--;
            __syncthreads();
        }
        // the result set in beam order, compacted in place, then rows by (distance, label)
        int got = 0;
        for (int b = 0; b < n; b += 64) {
            const int i = b + lane;
            float dv = 0.f;
            uint32_t iv = 0u;
            if (i < n) {
                dv = bd[i];
                iv = bi[i];
            }
            const bool al = i < n && (iv & DFLAG) == 0u;
            const uint64_t am = ballot(al);
            const int slot = got + lanes_below(am, lane);
            __asm__ volatile("" ::: "memory");
            if (al) {
                bd[slot] = dv;
                bi[slot] = iv;
            }
            __asm__ volatile("" ::: "memory");
            got += __popcll(am);
        }
        __syncthreads();
        const int row_len = (int)P.k_stride;
        uint64_t* ol = P.out_labels + (size_t)qi * P.k_stride;
        float* od = P.out_dists + (size_t)qi * P.k_stride;
        for (int i = lane; i < got; i += 64) {
            const float di = bd[i];
            const uint64_t li = g.labels[bi[i] & ID_MASK];
            int s0 = i;
            while (s0 > 0 && bd[s0 - 1] == di) s0--;
            int r = s0;
            for (int j = s0; j < got && bd[j] == di; j++) {
                if (j == i) continue;
                const uint64_t lj = g.labels[bi[j] & ID_MASK];
                r += (lj < li || (lj == li && j < i)) ? 1 : 0;
            }
            if (r < row_len) {
                ol[r] = li;
                od[r] = di;
            }
        }
        for (int i = (got < row_len ? got : row_len) + lane; i < row_len; i += 64) {
            ol[i] = ~0ull;
            od[i] = INFINITY;
        }
        if (lane == 0) {
            P.counts[qi] = (uint32_t)got;
            if (ovf) atomicOr(P.status, 1u);
            if (P.counters) {
                int32_t* c = P.counters + (size_t)qi * 4;
                c[0] = H0;
                c[1] = E0;
                c[2] = Hu;
                c[3] = Eu;
            }
        }
        __syncthreads();
    }
}

size_t mv_lds_bytes(uint32_t dim, uint32_t cap, uint32_t hash_bits, uint32_t dcap) {
    const size_t dimpad = (dim + 3u) & ~3u;
    return dimpad * 4 + (size_t)cap * 8 + ((size_t)1 << hash_bits) * 4 + 64 * 8 + (size_t)dcap * 8;
}

template <int SPACE, int MODE, int NCH>
static hipError_t launch_mv_t(const MvParams& p, int blocks, hipStream_t s) {
    const size_t lds = mv_lds_bytes(p.g.dim, p.cap, p.hash_bits, p.dcap);
    auto k = hnsw_mv_kernel<SPACE, MODE, NCH>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, dim3(blocks), dim3(64), lds, s, p);
    return hipGetLastError();
}

template <int SPACE>
static hipError_t launch_mv_dim(const MvParams& p, int blocks, hipStream_t s) {
    if (p.g.dim % 16 != 0) return launch_mv_t<SPACE, 0, 0>(p, blocks, s);
    if (p.g.dim == 128) return launch_mv_t<SPACE, 1, 8>(p, blocks, s);
    return launch_mv_t<SPACE, 1, 0>(p, blocks, s);
}

hipError_t launch_mv_search(const MvParams& p, int blocks, hipStream_t s) {
    hipError_t e = launch_zero_u32(p.work, 1, s);
    if (e != hipSuccess) return e;
    return p.g.space == 0 ? launch_mv_dim<0>(p, blocks, s) : launch_mv_dim<1>(p, blocks, s);
}

int mv_blocks_per_cu(const MvParams& p) {
    const size_t lds = mv_lds_bytes(p.g.dim, p.cap, p.hash_bits, p.dcap);
    int nb = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, hnsw_mv_kernel<0, 1, 0>, 64, lds);
    if (e != hipSuccess || nb <= 0) nb = 1;
    return nb;
}

}  // namespace hg
