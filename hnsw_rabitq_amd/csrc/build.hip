// Level-0 graph construction on the GPU, in batches of new elements.
//
// The reference inserts elements one at a time (addPoint, hnswalg.h:1152-1266): a beam search
// with ef_construction from the entry found by the upper-level descent (searchBaseLayer,
// hnswalg.h:225-305), neighbour selection by the heuristic (getNeighborsByHeuristic2,
// hnswalg.h:443-483), then mutual linking, where a full neighbour list is re-selected by the same
// heuristic (mutuallyConnectNewElement, hnswalg.h:506-630).  Here a batch of elements is inserted
// against the graph of the elements before it: the beam search is the query kernel, and two
// kernels of this file complete the candidates and link:
//   build_select_kernel  one wave per new element: merges into its candidate list the earlier
//                        elements of its own batch (brute force over them, so that every element
//                        still sees all elements inserted before it), then selects at most M
//                        neighbours by the heuristic and writes its list;
//   build_link_kernel    one wave per element that a new element selected: appends the new
//                        elements that selected it, or, when its list would overflow, re-selects
//                        maxM0 neighbours by the heuristic from its list and the new elements.
// Distances are the reference's FP32 values (the query kernel's distance routines); equal
// distances order by element id.  The result is a valid HNSW level-0 graph of the same kind as
// the reference's, not the same graph: insertion order within a batch differs.
#include <hipcub/hipcub.hpp>

#include "common.h"
#include "dist.h"
#include "kernels.h"
#include "wave.h"

namespace hg {

namespace {

__device__ __forceinline__ void load_vec(const DevGraph& g, uint32_t id, float* out, uint32_t dimpad, int lane) {
    const float* x = g.vec + (size_t)id * g.stride;
    for (uint32_t i = lane; i < dimpad; i += 64) out[i] = i < g.dim ? x[i] : 0.f;
}

// Heuristic selection (getNeighborsByHeuristic2): walk the candidates (cd, ci)[0..nc) in
// ascending (distance, id) order and keep one when no kept neighbour is closer to it than the
// query is, until `limit` are kept.  Kept ids go to S; returns their number.  q2 is scratch for
// the candidate vector, sd for distances (>= limit floats).
template <int SPACE, int MODE, int NCH>
__device__ int select_heuristic(const DevGraph& g, const float* cd, const uint32_t* ci, int nc, int limit,
                                uint32_t* S, float* q2, float* sd, uint32_t dimpad, int lane) {
    int nS = 0;
    if (nc < limit) {  // fewer candidates than the bound: all are kept (hnswalg.h:447-448)
        for (int i = lane; i < nc; i += 64) S[i] = ci[i];
        __syncthreads();
        return nc;
    }
    for (int c = 0; c < nc && nS < limit; c++) {
        const uint32_t cid = ci[c];
        const float dc = cd[c];
        bool good = true;
        if (nS > 0) {
            load_vec(g, cid, q2, dimpad, lane);
            __syncthreads();
            dist_batch<SPACE, MODE, NCH>(g, q2, S, nS, sd, lane);
            __syncthreads();
            good = ballot(lane < nS && sd[lane] < dc) == 0ull;
        }
        if (good) {
            if (lane == 0) S[nS] = cid;
            nS++;
            __syncthreads();
        }
    }
    return nS;
}

}  // namespace

template <int SPACE, int MODE, int NCH>
__global__ __launch_bounds__(64) void build_select_kernel(BuildArgs A) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const DevGraph& g = A.g;
    const uint32_t dimpad = (g.dim + 3u) & ~3u;
    const int efc = (int)A.efc;
    const int capL = (efc + 63) & ~63;
    float* qs = reinterpret_cast<float*>(smem);
    float* q2 = qs + dimpad;
    float* Ld = q2 + dimpad;
    uint32_t* Li = reinterpret_cast<uint32_t*>(Ld + capL);
    float* sd = reinterpret_cast<float*>(Li + capL);
    uint32_t* si = reinterpret_cast<uint32_t*>(sd + 64);
    uint32_t* S = si + 64;

    const uint32_t bx = blockIdx.x;
    const uint32_t p = A.s + bx;
    load_vec(g, p, qs, dimpad, lane);
    // candidates of the graph search, ascending, padded with ~0 ids
    int nL = 0;
    for (int b = 0; b < capL; b += 64) {
        const int i = b + lane;
        uint64_t id = ~0ull;
        float d = INFINITY;
        if (i < efc) {
            id = A.cand_ids[(size_t)bx * efc + i];
            d = A.cand_d[(size_t)bx * efc + i];
        }
        const bool v = id != ~0ull;
        Ld[i] = d;
        Li[i] = v ? (uint32_t)id : SENT;
        nL += __popcll(ballot(v));
    }
    __syncthreads();

    // earlier elements of the batch, 64 at a time
    for (uint32_t b = A.s; b < p; b += 64) {
        const int cnt = (int)((p - b) < 64u ? (p - b) : 64u);
        if (lane < cnt) si[lane] = b + lane;
        __syncthreads();
        dist_batch<SPACE, MODE, NCH>(g, qs, si, cnt, sd, lane);
        __syncthreads();
        const float d = lane < cnt ? sd[lane] : INFINITY;
        const uint32_t id = lane < cnt ? si[lane] : SENT;
        uint64_t m = ballot(lane < cnt && (nL < efc || key_less(d, id, Ld[nL - 1], Li[nL - 1])));
        while (m) {
            const int L = __ffsll((unsigned long long)m) - 1;
            m &= m - 1;
            const float dl = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(d), L));
            const uint32_t il = (uint32_t)__builtin_amdgcn_readlane((int)id, L);
            if (nL == efc && !key_less(dl, il, Ld[nL - 1], Li[nL - 1])) continue;
            int pos = 0;
            for (int c = 0; c < nL; c += 64) {
                const int i = c + lane;
                pos += __popcll(ballot(i < nL && key_less(Ld[i], Li[i], dl, il)));
            }
            const int newn = nL < efc ? nL + 1 : efc;
            // shift [pos, newn - 1) up by one, top chunk first; a chunk's reads precede its
            // writes and a wave's LDS operations execute in order
            for (int c = (newn - 1) & ~63; c >= (pos & ~63); c -= 64) {
                const int i = c + lane;
                const bool mv = i > pos && i < newn;
                float vd = 0.f;
                uint32_t vi = 0u;
                if (mv) {
                    vd = Ld[i - 1];
                    vi = Li[i - 1];
                }
                __asm__ volatile("" ::: "memory");
                if (mv) {
                    Ld[i] = vd;
                    Li[i] = vi;
                }
                __asm__ volatile("" ::: "memory");
            }
            if (lane == 0) {
                Ld[pos] = dl;
                Li[pos] = il;
            }
            nL = newn;
            __syncthreads();
        }
    }

    const int M = (int)A.M;
    const int nS = select_heuristic<SPACE, MODE, NCH>(g, Ld, Li, nL, M, S, q2, sd, dimpad, lane);
    // the new element's list, farthest first as mutuallyConnectNewElement writes it
    const uint32_t maxM0 = g.maxM0;
    uint32_t* row = A.links0 + (size_t)p * maxM0;
    for (uint32_t j = lane; j < maxM0; j += 64) row[j] = (int)j < nS ? S[nS - 1 - (int)j] : SENT;
    for (int j = lane; j < M; j += 64) A.sel[(size_t)bx * M + j] = j < nS ? S[j] : SENT;
}

template <int SPACE, int MODE, int NCH>
__global__ __launch_bounds__(64) void build_link_kernel(BuildArgs A) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const DevGraph& g = A.g;
    const uint32_t dimpad = (g.dim + 3u) & ~3u;
    constexpr int UCAP = kBuildUnionCap;
    float* qs = reinterpret_cast<float*>(smem);
    float* q2 = qs + dimpad;
    float* Ud = q2 + dimpad;
    uint32_t* Ui = reinterpret_cast<uint32_t*>(Ud + UCAP);
    float* Td = reinterpret_cast<float*>(Ui + UCAP);
    uint32_t* Ti = reinterpret_cast<uint32_t*>(Td + UCAP);
    uint32_t* cur = Ti + UCAP;        // [64] the element's list
    uint32_t* S = cur + 64;           // [64] selection
    float* sd = reinterpret_cast<float*>(S + 64);  // [64]

    const uint32_t t = blockIdx.x;
    const uint32_t T = *A.n_targets;
    if (t >= T) return;
    const uint32_t b0 = A.starts[t], b1 = t + 1 < T ? A.starts[t + 1] : *A.n_links;
    const uint32_t n = A.keys[b0];
    const int maxM0 = (int)g.maxM0;
    uint32_t* row = A.links0 + (size_t)n * maxM0;
    const uint32_t v = lane < maxM0 ? row[lane] : SENT;
    int c = __popcll(ballot(v != SENT));
    cur[lane] = v;
    load_vec(g, n, qs, dimpad, lane);
    __syncthreads();
    const int G = UCAP - maxM0;
    for (uint32_t gb = b0; gb < b1; gb += (uint32_t)G) {
        const int gn = (int)((b1 - gb) < (uint32_t)G ? (b1 - gb) : (uint32_t)G);
        if (c + gn <= maxM0) {
            if (lane < gn) cur[c + lane] = A.incoming[gb + lane];
            c += gn;
            __syncthreads();
            continue;
        }
        // list and arrivals together, by distance to the element, then the heuristic
        const int nU = c + gn;
        for (int i = lane; i < nU; i += 64) Ui[i] = i < c ? cur[i] : A.incoming[gb + (i - c)];
        __syncthreads();
        dist_batch<SPACE, MODE, NCH>(g, qs, Ui, nU, Ud, lane);
        __syncthreads();
        for (int i = lane; i < nU; i += 64) {
            const float di = Ud[i];
            const uint32_t ii = Ui[i];
            int r = 0;
            for (int j = 0; j < nU; j++) r += key_less(Ud[j], Ui[j], di, ii) ? 1 : 0;
            Td[r] = di;
            Ti[r] = ii;
        }
        __syncthreads();
        const int nS = select_heuristic<SPACE, MODE, NCH>(g, Td, Ti, nU, maxM0, S, q2, sd, dimpad, lane);
        // farthest first, as the reference rewrites a pruned list
        if (lane < nS) cur[lane] = S[nS - 1 - lane];
        c = nS;
        __syncthreads();
    }
    if (lane < maxM0) row[lane] = lane < c ? cur[lane] : SENT;
}

namespace {

// Link pairs of the batch: key = receiving element, value = new element (s + i).  Slots after a
// row's first SENT are unused and take the key s + B, above every receiver (receivers of row i
// are elements < s + i), so they sort last.
__global__ void link_pairs_kernel(const uint32_t* __restrict__ sel, uint32_t B, uint32_t M, uint32_t s,
                                  uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, uint32_t* err) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * M) return;
    const uint32_t i = idx / M, j = idx % M;
    bool ok = true;
    for (uint32_t q = 0; q <= j; q++) ok = ok && sel[i * M + q] != SENT;
    const uint32_t t = sel[idx];
    if (ok && t >= s + i) {  // not an earlier element: reported, and never used as a receiver
        atomicOr(err, 1u);
        ok = false;
    }
    keys[idx] = ok ? t : s + B;
    vals[idx] = s + i;
}

// After the sort: flags[idx] = 1 at the first link of each receiver; *n_links = number of links.
__global__ void link_runs_kernel(const uint32_t* __restrict__ keys, uint32_t N, uint32_t key_unused,
                                 uint32_t* __restrict__ flags, uint32_t* n_links) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    const uint32_t k = keys[idx];
    const bool valid = k < key_unused;
    flags[idx] = (valid && (idx == 0 || keys[idx - 1] != k)) ? 1u : 0u;
    if (valid && (idx + 1 == N || keys[idx + 1] >= key_unused)) *n_links = idx + 1;
}

__global__ void zero_words_kernel(uint32_t* p, uint32_t n) {
    if (threadIdx.x < n) p[threadIdx.x] = 0u;
}

struct GroupScratch {
    uint32_t *keys_in, *vals_in, *keys, *vals, *flags, *starts, *words;  // words: n_targets, n_links
    void* temp;
    size_t temp_bytes;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t cub_temp_bytes(size_t max_links) {
    size_t a = 0, b = 0;
    uint32_t* np = nullptr;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, a, np, np, np, np, (int)max_links, 0, 32);
    (void)hipcub::DeviceSelect::Flagged(nullptr, b, hipcub::CountingInputIterator<uint32_t>(0), np, np, np,
                                        (int)max_links);
    return a > b ? a : b;
}

GroupScratch carve(void* base, size_t max_links) {
    GroupScratch g;
    char* p = (char*)base;
    auto take = [&](size_t bytes) {
        char* r = p;
        p += align256(bytes);
        return r;
    };
    g.keys_in = (uint32_t*)take(max_links * 4);
    g.vals_in = (uint32_t*)take(max_links * 4);
    g.keys = (uint32_t*)take(max_links * 4);
    g.vals = (uint32_t*)take(max_links * 4);
    g.flags = (uint32_t*)take(max_links * 4);
    g.starts = (uint32_t*)take(max_links * 4);
    g.words = (uint32_t*)take(16);
    g.temp_bytes = cub_temp_bytes(max_links);
    g.temp = take(g.temp_bytes);
    return g;
}
}  // namespace

size_t build_group_scratch_bytes(size_t max_links) {
    return 6 * align256(max_links * 4) + align256(16) + align256(cub_temp_bytes(max_links)) + 256;
}

hipError_t build_group_links(BuildArgs& a, void* scratch, size_t scratch_bytes, uint32_t* err, hipStream_t s) {
    const size_t N = (size_t)a.B * a.M;
    if (N == 0 || build_group_scratch_bytes(N) > scratch_bytes) return hipErrorInvalidValue;
    GroupScratch g = carve(scratch, N);
    const uint32_t key_unused = a.s + a.B;
    int end_bit = 1;
    while (end_bit < 32 && (1ull << end_bit) <= key_unused) end_bit++;
    const uint32_t blocks = (uint32_t)((N + 255) / 256);
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, s, g.words, 4u);
    hipLaunchKernelGGL(link_pairs_kernel, dim3(blocks), dim3(256), 0, s, a.sel, a.B, a.M, a.s, g.keys_in, g.vals_in,
                       err);
    size_t tb = g.temp_bytes;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(g.temp, tb, g.keys_in, g.keys, g.vals_in, g.vals, (int)N, 0,
                                                      end_bit, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(link_runs_kernel, dim3(blocks), dim3(256), 0, s, g.keys, (uint32_t)N, key_unused, g.flags,
                       g.words + 1);
    tb = g.temp_bytes;
    e = hipcub::DeviceSelect::Flagged(g.temp, tb, hipcub::CountingInputIterator<uint32_t>(0), g.flags, g.starts,
                                      g.words, (int)N, s);
    if (e != hipSuccess) return e;
    a.keys = g.keys;
    a.incoming = g.vals;
    a.starts = g.starts;
    a.n_targets = g.words;
    a.n_links = g.words + 1;
    return hipGetLastError();
}

size_t build_select_lds(uint32_t dim, uint32_t efc) {
    const size_t dimpad = (dim + 3u) & ~3u;
    const size_t capL = (efc + 63u) & ~63u;
    return dimpad * 8 + capL * 8 + 64 * 12;
}

size_t build_link_lds(uint32_t dim) {
    const size_t dimpad = (dim + 3u) & ~3u;
    return dimpad * 8 + (size_t)kBuildUnionCap * 16 + 64 * 12;
}

template <int SPACE, int MODE, int NCH>
static hipError_t launch_build_t(const BuildArgs& a, bool link, hipStream_t s) {
    const size_t lds = link ? build_link_lds(a.g.dim) : build_select_lds(a.g.dim, a.efc);
    auto k = link ? build_link_kernel<SPACE, MODE, NCH> : build_select_kernel<SPACE, MODE, NCH>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const uint32_t grid = link ? a.B * a.M : a.B;  // link: receivers beyond *n_targets exit
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(k, dim3(grid), dim3(64), lds, s, a);
    return hipGetLastError();
}

template <int SPACE>
static hipError_t launch_build_dim(const BuildArgs& a, bool link, hipStream_t s) {
    if (a.g.dim % 16 != 0) return launch_build_t<SPACE, 0, 0>(a, link, s);
    switch (a.g.dim / 16) {
        case 8: return launch_build_t<SPACE, 1, 8>(a, link, s);
        default: return launch_build_t<SPACE, 1, 0>(a, link, s);
    }
}

hipError_t launch_build(const BuildArgs& a, bool link, hipStream_t s) {
    return a.g.space == 0 ? launch_build_dim<0>(a, link, s) : launch_build_dim<1>(a, link, s);
}

}  // namespace hg
