// CPU oracle (test infrastructure only; see oracle.h for the function-to-reference map).
// Compiled with -ffp-contract=off: every fused multiply-add below is an explicit fmaf, so the
// arithmetic order and rounding are exactly the ones written here, on any x86-64 host.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <fstream>
#include <functional>
#include <limits>
#include <mutex>
#include <queue>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_err;

// ---------------------------------------------------------------------------------------------
// Distances.  Kernel selection follows L2Space / InnerProductSpace constructors
// (space_l2.h:214-238, space_ip.h:348-383) on a host where AVX512Capable() and AVXCapable()
// are true: SIMD16 -> AVX-512 kernel, SIMD4 (L2) -> SSE kernel, SIMD4 (IP) -> AVX kernel.
// ---------------------------------------------------------------------------------------------

// Scalar L2 loop (space_l2.h:6-20) as GCC -O3 -march=native compiles it: groups of four
// squares are formed with one vector multiply (each product rounded) and added to the running
// sum in order; the remaining n % 4 terms use one fma each.  The IP loop (space_ip.h:7-14) is
// one fma per term.  Both pinned by the distance golden vectors at every dimension class.
inline float l2_scalar(const float* a, const float* b, size_t n) {
    float res = 0.f;
    size_t nv = n / 4 * 4;
    for (size_t i = 0; i < nv; i++) {
        float t = a[i] - b[i];
        float sq = t * t;
        res = res + sq;
    }
    for (size_t i = nv; i < n; i++) {
        float t = a[i] - b[i];
        res = fmaf(t, t, res);
    }
    return res;
}
inline float ip_scalar(const float* a, const float* b, size_t n) {
    float res = 0.f;
    for (size_t i = 0; i < n; i++) res = fmaf(a[i], b[i], res);
    return res;
}

// L2SqrSIMD16ExtAVX512 (space_l2.h:25-54): 16 lane accumulators acc[l] = fma(d,d,acc[l]) over
// chunks of 16, then the 16 lanes summed left to right.
inline float l2_simd16(const float* a, const float* b, size_t n16) {
    float acc[16] = {0};
    for (size_t c = 0; c < n16; c += 16)
        for (int l = 0; l < 16; l++) {
            float d = a[c + l] - b[c + l];
            acc[l] = fmaf(d, d, acc[l]);
        }
    float r = acc[0];
    for (int l = 1; l < 16; l++) r = r + acc[l];
    return r;
}

// L2SqrSIMD4Ext (SSE, space_l2.h:165-186): 4 lane accumulators, lanes summed left to right.
inline float l2_simd4(const float* a, const float* b, size_t n4) {
    float acc[4] = {0};
    for (size_t c = 0; c < n4; c += 4)
        for (int l = 0; l < 4; l++) {
            float d = a[c + l] - b[c + l];
            acc[l] = fmaf(d, d, acc[l]);
        }
    return ((acc[0] + acc[1]) + acc[2]) + acc[3];
}

// InnerProductSIMD16ExtAVX512 (space_ip.h:146-199): fma into 16 lanes, then GCC's
// _mm512_reduce_add_ps: halves (i, i+8), then (i, i+4), then (i, i+2), then the last pair.
inline float ip_simd16(const float* a, const float* b, size_t n16) {
    float acc[16] = {0};
    for (size_t c = 0; c < n16; c += 16)
        for (int l = 0; l < 16; l++) acc[l] = fmaf(a[c + l], b[c + l], acc[l]);
    float h8[8], h4[4], h2[2];
    for (int i = 0; i < 8; i++) h8[i] = acc[i] + acc[i + 8];
    for (int i = 0; i < 4; i++) h4[i] = h8[i] + h8[i + 4];
    for (int i = 0; i < 2; i++) h2[i] = h4[i] + h4[i + 2];
    return h2[0] + h2[1];
}

// InnerProductSIMD4ExtAVX (space_ip.h:24-69): 8 lanes over chunks of 16 (two 8-wide fmas per
// chunk), fold high half onto low half, 4-wide fmas over the remaining chunks of 4, then the 4
// lanes summed left to right.
inline float ip_simd4_avx(const float* a, const float* b, size_t n) {
    size_t n16 = n / 16 * 16, n4 = n / 4 * 4;
    float acc8[8] = {0};
    for (size_t c = 0; c < n16; c += 8)
        for (int l = 0; l < 8; l++) acc8[l] = fmaf(a[c + l], b[c + l], acc8[l]);
    float acc[4];
    for (int l = 0; l < 4; l++) acc[l] = acc8[l] + acc8[l + 4];
    for (size_t c = n16; c < n4; c += 4)
        for (int l = 0; l < 4; l++) acc[l] = fmaf(a[c + l], b[c + l], acc[l]);
    return ((acc[0] + acc[1]) + acc[2]) + acc[3];
}

float dist_l2(size_t dim, const float* a, const float* b) {
    if (dim % 16 == 0) return l2_simd16(a, b, dim);
    if (dim % 4 == 0) return l2_simd4(a, b, dim);
    if (dim > 16) {  // L2SqrSIMD16ExtResiduals, space_l2.h:149-160
        size_t n16 = dim >> 4 << 4;
        float res = l2_simd16(a, b, n16);
        float tail = l2_scalar(a + n16, b + n16, dim - n16);
        return res + tail;
    }
    if (dim > 4) {  // L2SqrSIMD4ExtResiduals, space_l2.h:189-205
        size_t n4 = dim >> 2 << 2;
        float res = l2_simd4(a, b, n4);
        float tail = l2_scalar(a + n4, b + n4, dim - n4);
        return res + tail;
    }
    return l2_scalar(a, b, dim);
}

float dist_ip(size_t dim, const float* a, const float* b) {
    if (dim % 16 == 0) return 1.0f - ip_simd16(a, b, dim);
    if (dim % 4 == 0) return 1.0f - ip_simd4_avx(a, b, dim);
    if (dim > 16) {  // InnerProductDistanceSIMD16ExtResiduals, space_ip.h:307-318
        size_t n16 = dim >> 4 << 4;
        float res = ip_simd16(a, b, n16);
        float tail = ip_scalar(a + n16, b + n16, dim - n16);
        return 1.0f - (res + tail);
    }
    if (dim > 4) {  // InnerProductDistanceSIMD4ExtResiduals, space_ip.h:320-332
        size_t n4 = dim >> 2 << 2;
        float res = ip_simd4_avx(a, b, n4);
        float tail = ip_scalar(a + n4, b + n4, dim - n4);
        return 1.0f - (res + tail);
    }
    return 1.0f - ip_scalar(a, b, dim);  // InnerProductDistance, space_ip.h:16-19
}

inline float dist(int space, size_t dim, const float* a, const float* b) {
    return space == OR_SPACE_L2 ? dist_l2(dim, a, b) : dist_ip(dim, a, b);
}

// ParallelFor over rows with an atomic counter (bindings.cpp:22-74).
void parallel_for(size_t n, int nthreads, const std::function<void(size_t, int)>& fn) {
    if (nthreads <= 0) nthreads = (int)std::thread::hardware_concurrency();
    if (nthreads <= 1 || n <= 1) {
        for (size_t i = 0; i < n; i++) fn(i, 0);
        return;
    }
    std::atomic<size_t> next(0);
    std::vector<std::thread> ts;
    for (int t = 0; t < nthreads; t++)
        ts.emplace_back([&, t] {
            for (;;) {
                size_t i = next.fetch_add(1);
                if (i >= n) break;
                fn(i, t);
            }
        });
    for (auto& th : ts) th.join();
}

// ---------------------------------------------------------------------------------------------
// HNSW index as stored on disk (hnswalg.h:685-712) and held in memory (hnswalg.h:120-143).
// ---------------------------------------------------------------------------------------------
struct Hnsw {
    size_t offsetLevel0 = 0, max_elements = 0, count = 0, size_per_element = 0, label_offset = 0,
           offsetData = 0;
    int maxlevel = 0;
    uint32_t entry = 0;
    size_t maxM = 0, maxM0 = 0, M = 0;
    double mult = 0;
    size_t efc = 0;
    size_t size_links_per_element = 0;
    std::vector<char> level0;
    std::vector<std::vector<char>> links;
    std::vector<int> levels;
    size_t dim = 0;

    const uint32_t* ll0(uint32_t i) const {
        return (const uint32_t*)(level0.data() + i * size_per_element + offsetLevel0);
    }
    const uint32_t* ll(uint32_t i, int level) const {
        return (const uint32_t*)(links[i].data() + (level - 1) * size_links_per_element);
    }
    const float* vec(uint32_t i) const { return (const float*)(level0.data() + i * size_per_element + offsetData); }
    uint64_t label(uint32_t i) const {
        uint64_t l;
        memcpy(&l, level0.data() + i * size_per_element + label_offset, 8);
        return l;
    }
    static unsigned short count_of(const uint32_t* p) { return *(const unsigned short*)p; }
    // isMarkedDeleted (hnswalg.h:930-935): bit 0 of byte 2 of the level-0 header
    bool deleted(uint32_t i) const { return (((const unsigned char*)ll0(i))[2] & 0x01) != 0; }
    size_t num_deleted = 0;
};

template <typename T>
bool rd(const char*& p, const char* end, T& v) {
    if (p + sizeof(T) > end) return false;
    memcpy(&v, p, sizeof(T));
    p += sizeof(T);
    return true;
}

Hnsw* parse(const char* buf, size_t len) {
    const char* p = buf;
    const char* end = buf + len;
    Hnsw* h = new Hnsw();
    bool ok = rd(p, end, h->offsetLevel0) && rd(p, end, h->max_elements) && rd(p, end, h->count) &&
              rd(p, end, h->size_per_element) && rd(p, end, h->label_offset) && rd(p, end, h->offsetData) &&
              rd(p, end, h->maxlevel) && rd(p, end, h->entry) && rd(p, end, h->maxM) && rd(p, end, h->maxM0) &&
              rd(p, end, h->M) && rd(p, end, h->mult) && rd(p, end, h->efc);
    // the record layout of hnswalg.h:121-127 (links, vector, label) and sizes that fit the file,
    // with the products checked for overflow
    ok = ok && h->size_per_element > 0 && h->count <= (size_t)(end - p) / h->size_per_element &&
         h->offsetLevel0 == 0 && h->maxM0 > 0 && h->maxM > 0 && h->maxM0 < (1u << 20) && h->maxM < (1u << 20) &&
         h->offsetData == h->maxM0 * 4 + 4 && h->label_offset > h->offsetData &&
         (h->label_offset - h->offsetData) % 4 == 0 && h->size_per_element == h->label_offset + 8 &&
         h->max_elements >= h->count && (h->count == 0 || h->entry < h->count) && h->maxlevel >= 0;
    if (!ok) {
        delete h;
        throw std::runtime_error("Index seems to be corrupted or unsupported");
    }
    h->level0.assign(p, p + h->count * h->size_per_element);
    p += h->count * h->size_per_element;
    h->size_links_per_element = h->maxM * 4 + 4;
    h->links.resize(h->count);
    h->levels.assign(h->count, 0);
    for (size_t i = 0; i < h->count; i++) {
        uint32_t sz;
        if (!rd(p, end, sz) || (size_t)(end - p) < sz) {
            delete h;
            throw std::runtime_error("Index seems to be corrupted or unsupported");
        }
        if (sz % h->size_links_per_element != 0) {
            delete h;
            throw std::runtime_error("Index seems to be corrupted or unsupported");
        }
        if (sz) {
            h->links[i].assign(p, p + sz);
            h->levels[i] = (int)(sz / h->size_links_per_element);
            p += sz;
        }
    }
    if (p != end) {
        delete h;
        throw std::runtime_error("Index seems to be corrupted or unsupported");
    }
    h->dim = (h->label_offset - h->offsetData) / 4;
    // every list within its capacity and every id an element; the entry point holds the top level
    bool lists_ok = h->count == 0 || h->levels[h->entry] >= h->maxlevel;
    for (size_t i = 0; lists_ok && i < h->count; i++) {
        const uint32_t* l0 = reinterpret_cast<const uint32_t*>(h->level0.data() + i * h->size_per_element);
        uint16_t c0;
        memcpy(&c0, l0, 2);
        lists_ok = c0 <= h->maxM0;
        for (size_t j = 0; lists_ok && j < c0; j++) {
            uint32_t id;
            memcpy(&id, l0 + 1 + j, 4);
            lists_ok = id < h->count;
        }
        for (int lv = 1; lists_ok && lv <= h->levels[i]; lv++) {
            const char* l = h->links[i].data() + (size_t)(lv - 1) * h->size_links_per_element;
            uint16_t c;
            memcpy(&c, l, 2);
            lists_ok = c <= h->maxM;
            for (size_t j = 0; lists_ok && j < c; j++) {
                uint32_t id;
                memcpy(&id, l + 4 + 4 * j, 4);
                lists_ok = id < h->count;
            }
        }
    }
    if (!lists_ok) {
        delete h;
        throw std::runtime_error("Index seems to be corrupted or unsupported");
    }
    for (size_t i = 0; i < h->count; i++) h->num_deleted += h->deleted((uint32_t)i) ? 1 : 0;
    return h;
}

struct Res {
    float d;
    uint32_t id;
};

// One query, reference searchKnn (hnswalg.h:1269-1323) with searchBaseLayerST (hnswalg.h:308-440):
// the bare-bone branch when the index has no deleted element and no filter is given, else the
// non-bare-bone branch (deleted or filtered nodes are traversed but never enter the result set;
// the search stops only once the result set holds ef elements).  Mode HEAP uses
// std::priority_queue with CompareByFirst; mode TOTAL keys every comparison by (dist, id).
struct Searcher {
    const Hnsw& h;
    int space;
    std::vector<uint16_t> visited;
    uint16_t tag = 0;
    Searcher(const Hnsw& hh, int s) : h(hh), space(s), visited(hh.count, 0) {}

    void next_tag() {
        tag++;
        if (tag == 0) {  // VisitedList::reset, visited_list_pool.h:22-28
            std::fill(visited.begin(), visited.end(), 0);
            tag = 1;
        }
    }

    typedef std::pair<float, uint32_t> P;
    struct ByFirst {
        bool operator()(const P& a, const P& b) const { return a.first < b.first; }
    };
    struct ByPair {
        bool operator()(const P& a, const P& b) const { return a < b; }
    };
    struct ByPairNeg {  // min-heap on (dist,id) held as max-heap
        bool operator()(const P& a, const P& b) const { return b < a; }
    };

    // hop_cap > 0: the bare-bone search also stops once it has expanded hop_cap level-0 nodes
    // (a stop condition whose should_stop_search fires after that many hops; hnswlib.h:134-150,
    // hnswalg.h:343-355)
    template <int MODE>
    void run(const float* q, size_t k, size_t ef, std::vector<Res>& out, int64_t* ctr, uint8_t* tie,
             const uint8_t* allowed = nullptr, int64_t hop_cap = 0) {
        out.clear();
        int64_t H0 = 0, E0 = 0, Hu = 0, Eu = 0;
        if (h.count == 0) return;
        size_t dim = h.dim;
        uint32_t cur = h.entry;
        float curdist = dist(space, dim, q, h.vec(cur));
        Eu++;
        for (int level = h.maxlevel; level > 0; level--) {
            bool changed = true;
            while (changed) {
                changed = false;
                const uint32_t* data = h.ll(cur, level);
                int size = Hnsw::count_of(data);
                Hu++;
                const uint32_t* datal = data + 1;
                for (int i = 0; i < size; i++) {
                    uint32_t cand = datal[i];
                    if (cand > h.max_elements) throw std::runtime_error("cand error");
                    float d = dist(space, dim, q, h.vec(cand));
                    Eu++;
                    if (d < curdist) {
                        curdist = d;
                        cur = cand;
                        changed = true;
                    }
                }
            }
        }
        ef = std::max(ef, k);
        next_tag();
        std::vector<float> seen;  // every level-0 distance (tie detection)
        bool track = tie != nullptr;

        const bool bare = h.num_deleted == 0 && allowed == nullptr;
        auto ok = [&](uint32_t id) { return !h.deleted(id) && (allowed == nullptr || allowed[id] != 0); };
        if (!bare && MODE == OR_MODE_HEAP) {
            // searchBaseLayerST<false> (hnswalg.h:323-437)
            std::priority_queue<P, std::vector<P>, ByFirst> top, cand;
            float lowerBound;
            if (ok(cur)) {
                float dep = dist(space, dim, q, h.vec(cur));
                E0++;
                if (track) seen.push_back(dep);
                lowerBound = dep;
                top.emplace(dep, cur);
                cand.emplace(-dep, cur);
            } else {
                lowerBound = std::numeric_limits<float>::max();
                cand.emplace(-lowerBound, cur);
            }
            visited[cur] = tag;
            while (!cand.empty()) {
                P cp = cand.top();
                if (-cp.first > lowerBound && top.size() == ef) break;
                cand.pop();
                const uint32_t* data = h.ll0(cp.second);
                size_t size = Hnsw::count_of(data);
                H0++;
                for (size_t j = 1; j <= size; j++) {
                    uint32_t cid = data[j];
                    if (visited[cid] == tag) continue;
                    visited[cid] = tag;
                    float d = dist(space, dim, q, h.vec(cid));
                    E0++;
                    if (track) seen.push_back(d);
                    if (top.size() < ef || lowerBound > d) {
                        cand.emplace(-d, cid);
                        if (ok(cid)) top.emplace(d, cid);
                        while (top.size() > ef) top.pop();
                        if (!top.empty()) lowerBound = top.top().first;
                    }
                }
            }
            while (top.size() > k) top.pop();
            while (!top.empty()) {
                out.push_back({top.top().first, top.top().second});
                top.pop();
            }
        } else if (!bare) {
            // the same branch in (dist, id) order: the entry point's distance is the descent's
            // (counted once), and a disallowed entry point is expanded first
            std::priority_queue<P, std::vector<P>, ByPair> top;
            std::priority_queue<P, std::vector<P>, ByPairNeg> cand;
            E0++;
            if (track) seen.push_back(curdist);
            if (ok(cur)) top.emplace(curdist, cur);
            cand.emplace(-std::numeric_limits<float>::infinity(), cur);
            visited[cur] = tag;
            while (!cand.empty()) {
                P cp = cand.top();
                if (top.size() == ef && top.top() < cp) break;
                cand.pop();
                const uint32_t* data = h.ll0(cp.second);
                size_t size = Hnsw::count_of(data);
                H0++;
                for (size_t j = 1; j <= size; j++) {
                    uint32_t cid = data[j];
                    if (visited[cid] == tag) continue;
                    visited[cid] = tag;
                    float d = dist(space, dim, q, h.vec(cid));
                    E0++;
                    if (track) seen.push_back(d);
                    P np(d, cid);
                    if (top.size() < ef || np < top.top()) {
                        cand.push(np);
                        if (ok(cid)) top.push(np);
                        while (top.size() > ef) top.pop();
                    }
                }
            }
            while (top.size() > k) top.pop();
            while (!top.empty()) {
                out.push_back({top.top().first, top.top().second});
                top.pop();
            }
        } else if (MODE == OR_MODE_HEAP) {
            std::priority_queue<P, std::vector<P>, ByFirst> top, cand;
            float dep = dist(space, dim, q, h.vec(cur));
            E0++;
            if (track) seen.push_back(dep);
            float lowerBound = dep;
            top.emplace(dep, cur);
            cand.emplace(-dep, cur);
            visited[cur] = tag;
            while (!cand.empty()) {
                P cp = cand.top();
                if (-cp.first > lowerBound || (hop_cap > 0 && H0 >= hop_cap)) break;
                cand.pop();
                const uint32_t* data = h.ll0(cp.second);
                size_t size = Hnsw::count_of(data);
                H0++;
                for (size_t j = 1; j <= size; j++) {
                    uint32_t cid = data[j];
                    if (visited[cid] == tag) continue;
                    visited[cid] = tag;
                    float d = dist(space, dim, q, h.vec(cid));
                    E0++;
                    if (track) seen.push_back(d);
                    if (top.size() < ef || lowerBound > d) {
                        cand.emplace(-d, cid);
                        top.emplace(d, cid);
                        while (top.size() > ef) top.pop();
                        if (!top.empty()) lowerBound = top.top().first;
                    }
                }
            }
            while (top.size() > k) top.pop();
            while (!top.empty()) {
                out.push_back({top.top().first, top.top().second});
                top.pop();
            }
        } else {
            std::priority_queue<P, std::vector<P>, ByPair> top;
            std::priority_queue<P, std::vector<P>, ByPairNeg> cand;
            float dep = dist(space, dim, q, h.vec(cur));
            E0++;
            if (track) seen.push_back(dep);
            top.emplace(dep, cur);
            cand.emplace(dep, cur);
            visited[cur] = tag;
            while (!cand.empty()) {
                P cp = cand.top();
                if (top.top() < cp || (hop_cap > 0 && H0 >= hop_cap)) break;
                cand.pop();
                const uint32_t* data = h.ll0(cp.second);
                size_t size = Hnsw::count_of(data);
                H0++;
                for (size_t j = 1; j <= size; j++) {
                    uint32_t cid = data[j];
                    if (visited[cid] == tag) continue;
                    visited[cid] = tag;
                    float d = dist(space, dim, q, h.vec(cid));
                    E0++;
                    if (track) seen.push_back(d);
                    P np(d, cid);
                    if (top.size() < ef || np < top.top()) {
                        cand.push(np);
                        top.push(np);
                        while (top.size() > ef) top.pop();
                    }
                }
            }
            while (top.size() > k) top.pop();
            while (!top.empty()) {
                out.push_back({top.top().first, top.top().second});
                top.pop();
            }
        }
        if (ctr) {
            ctr[0] = H0;
            ctr[1] = E0;
            ctr[2] = Hu;
            ctr[3] = Eu;
        }
        if (tie) {
            std::sort(seen.begin(), seen.end());
            *tie = std::adjacent_find(seen.begin(), seen.end()) != seen.end() ? 1 : 0;
        }
    }

    // searchStopConditionClosest (hnswalg.h:1326-1377) with MultiVectorSearchStopCondition
    // (stop_condition.h:82-200): the result set is bounded by its number of distinct documents
    // (ef_collection = max(ef_collection, num_docs)); a result is evicted (farthest first) while
    // it holds more documents than that; the search stops at a candidate farther than the worst
    // result once the bound is met; the farthest results are dropped until num_docs documents
    // remain.  docs[internal id] = document.  `out` is closest first.
    template <int MODE>
    void run_mv(const float* q, const uint32_t* docs, size_t num_docs, size_t ef_coll, std::vector<Res>& out,
                int64_t* ctr, uint8_t* tie, const uint8_t* allowed) {
        out.clear();
        int64_t H0 = 0, E0 = 0, Hu = 0, Eu = 0;
        if (h.count == 0) return;
        const size_t dim = h.dim;
        const size_t efc = std::max(ef_coll, num_docs);
        uint32_t cur = h.entry;
        float curdist = dist(space, dim, q, h.vec(cur));
        Eu++;
        for (int level = h.maxlevel; level > 0; level--) {
            bool changed = true;
            while (changed) {
                changed = false;
                const uint32_t* data = h.ll(cur, level);
                int size = Hnsw::count_of(data);
                Hu++;
                for (int i = 0; i < size; i++) {
                    uint32_t cand = data[1 + i];
                    if (cand > h.max_elements) throw std::runtime_error("cand error");
                    float d = dist(space, dim, q, h.vec(cand));
                    Eu++;
                    if (d < curdist) {
                        curdist = d;
                        cur = cand;
                        changed = true;
                    }
                }
            }
        }
        next_tag();
        std::vector<float> seen;
        const bool track = tie != nullptr;
        auto ok = [&](uint32_t id) { return !h.deleted(id) && (allowed == nullptr || allowed[id] != 0); };
        const float fmax = std::numeric_limits<float>::max();
        std::unordered_map<uint32_t, size_t> counter;
        size_t ndocs = 0;
        auto add_doc = [&](uint32_t id) {
            if (counter[docs[id]]++ == 0) ndocs++;
        };
        auto remove_doc = [&](uint32_t doc) {
            if (--counter[doc] == 0) ndocs--;
        };
        std::vector<P> res;
        if (MODE == OR_MODE_HEAP) {
            std::priority_queue<std::pair<float, uint32_t>> sr;  // search_results_: (dist, doc)
            std::priority_queue<P, std::vector<P>, ByFirst> top, cand;
            float lowerBound;
            if (ok(cur)) {
                float dep = dist(space, dim, q, h.vec(cur));
                E0++;
                if (track) seen.push_back(dep);
                lowerBound = dep;
                top.emplace(dep, cur);
                add_doc(cur);
                sr.emplace(dep, docs[cur]);
                cand.emplace(-dep, cur);
            } else {
                lowerBound = fmax;
                cand.emplace(-lowerBound, cur);
            }
            visited[cur] = tag;
            while (!cand.empty()) {
                P cp = cand.top();
                if (-cp.first > lowerBound && ndocs == efc) break;
                cand.pop();
                const uint32_t* data = h.ll0(cp.second);
                size_t size = Hnsw::count_of(data);
                H0++;
                for (size_t j = 1; j <= size; j++) {
                    uint32_t cid = data[j];
                    if (visited[cid] == tag) continue;
                    visited[cid] = tag;
                    float d = dist(space, dim, q, h.vec(cid));
                    E0++;
                    if (track) seen.push_back(d);
                    if (ndocs < efc || lowerBound > d) {
                        cand.emplace(-d, cid);
                        if (ok(cid)) {
                            top.emplace(d, cid);
                            add_doc(cid);
                            sr.emplace(d, docs[cid]);
                        }
                        while (ndocs > efc) {  // remove_point_from_result pops search_results_'s top
                            remove_doc(docs[top.top().second]);
                            top.pop();
                            sr.pop();
                        }
                        if (!top.empty()) lowerBound = top.top().first;
                    }
                }
            }
            res.resize(top.size());
            for (size_t i = res.size(); i-- > 0;) {
                res[i] = top.top();
                top.pop();
            }
            while (ndocs > num_docs && !res.empty()) {  // filter_results: the doc from search_results_
                const uint32_t doc = sr.top().second;
                remove_doc(doc);
                sr.pop();
                res.pop_back();
            }
        } else {
            std::priority_queue<P, std::vector<P>, ByPair> top;
            std::priority_queue<P, std::vector<P>, ByPairNeg> cand;
            E0++;
            if (track) seen.push_back(curdist);
            if (ok(cur)) {
                top.emplace(curdist, cur);
                add_doc(cur);
            }
            cand.emplace(-std::numeric_limits<float>::infinity(), cur);
            visited[cur] = tag;
            bool first = true;
            while (!cand.empty()) {
                P cp = cand.top();
                const P key = first ? P(curdist, cur) : cp;
                first = false;
                if (ndocs == efc && !top.empty() && top.top() < key) break;
                cand.pop();
                const uint32_t* data = h.ll0(cp.second);
                size_t size = Hnsw::count_of(data);
                H0++;
                for (size_t j = 1; j <= size; j++) {
                    uint32_t cid = data[j];
                    if (visited[cid] == tag) continue;
                    visited[cid] = tag;
                    float d = dist(space, dim, q, h.vec(cid));
                    E0++;
                    if (track) seen.push_back(d);
                    P np(d, cid);
                    if (ndocs < efc || np < top.top()) {
                        cand.push(np);
                        if (ok(cid)) {
                            top.push(np);
                            add_doc(cid);
                        }
                        while (ndocs > efc) {
                            remove_doc(docs[top.top().second]);
                            top.pop();
                        }
                    }
                }
            }
            res.resize(top.size());
            for (size_t i = res.size(); i-- > 0;) {
                res[i] = top.top();
                top.pop();
            }
            while (ndocs > num_docs && !res.empty()) {
                remove_doc(docs[res.back().second]);
                res.pop_back();
            }
        }
        for (auto& r : res) out.push_back({r.first, r.second});
        if (ctr) {
            ctr[0] = H0;
            ctr[1] = E0;
            ctr[2] = Hu;
            ctr[3] = Eu;
        }
        if (tie) {
            std::sort(seen.begin(), seen.end());
            *tie = std::adjacent_find(seen.begin(), seen.end()) != seen.end() ? 1 : 0;
        }
    }

    // searchStopConditionClosest (hnswalg.h:1326-1377) with EpsilonSearchStopCondition
    // (stop_condition.h:207-275) driving searchBaseLayerST<false> (hnswalg.h:323-437): the result
    // set holds at most max_c allowed elements; the search stops at a candidate farther than the
    // result set's worst once it holds max_c elements, or at a candidate farther than eps once it
    // holds min_c; the results farther than eps are dropped.  `out` is closest first.
    template <int MODE>
    void run_eps(const float* q, float eps, size_t min_c, size_t max_c, std::vector<Res>& out, int64_t* ctr,
                 uint8_t* tie, const uint8_t* allowed) {
        out.clear();
        int64_t H0 = 0, E0 = 0, Hu = 0, Eu = 0;
        if (h.count == 0) return;
        const size_t dim = h.dim;
        uint32_t cur = h.entry;
        float curdist = dist(space, dim, q, h.vec(cur));
        Eu++;
        for (int level = h.maxlevel; level > 0; level--) {  // hnswalg.h:1335-1360
            bool changed = true;
            while (changed) {
                changed = false;
                const uint32_t* data = h.ll(cur, level);
                int size = Hnsw::count_of(data);
                Hu++;
                for (int i = 0; i < size; i++) {
                    uint32_t cand = data[1 + i];
                    if (cand > h.max_elements) throw std::runtime_error("cand error");
                    float d = dist(space, dim, q, h.vec(cand));
                    Eu++;
                    if (d < curdist) {
                        curdist = d;
                        cur = cand;
                        changed = true;
                    }
                }
            }
        }
        next_tag();
        std::vector<float> seen;
        const bool track = tie != nullptr;
        auto ok = [&](uint32_t id) { return !h.deleted(id) && (allowed == nullptr || allowed[id] != 0); };
        const float fmax = std::numeric_limits<float>::max();
        size_t curr = 0;  // curr_num_items_
        std::vector<P> res;
        if (MODE == OR_MODE_HEAP) {
            std::priority_queue<P, std::vector<P>, ByFirst> top, cand;
            float lowerBound;
            if (ok(cur)) {
                float dep = dist(space, dim, q, h.vec(cur));
                E0++;
                if (track) seen.push_back(dep);
                lowerBound = dep;
                top.emplace(dep, cur);
                curr++;
                cand.emplace(-dep, cur);
            } else {
                lowerBound = fmax;
                cand.emplace(-lowerBound, cur);
            }
            visited[cur] = tag;
            while (!cand.empty()) {
                P cp = cand.top();
                const float cd = -cp.first;
                if ((cd > lowerBound && curr == max_c) || (cd > eps && curr >= min_c)) break;
                cand.pop();
                const uint32_t* data = h.ll0(cp.second);
                size_t size = Hnsw::count_of(data);
                H0++;
                for (size_t j = 1; j <= size; j++) {
                    uint32_t cid = data[j];
                    if (visited[cid] == tag) continue;
                    visited[cid] = tag;
                    float d = dist(space, dim, q, h.vec(cid));
                    E0++;
                    if (track) seen.push_back(d);
                    if (curr < max_c || lowerBound > d) {
                        cand.emplace(-d, cid);
                        if (ok(cid)) {
                            top.emplace(d, cid);
                            curr++;
                        }
                        while (curr > max_c) {
                            top.pop();
                            curr--;
                        }
                        if (!top.empty()) lowerBound = top.top().first;
                    }
                }
            }
            res.resize(top.size());
            for (size_t i = res.size(); i-- > 0;) {
                res[i] = top.top();
                top.pop();
            }
        } else {
            // (dist, id) keys; a disallowed entry point is popped first with candidate distance
            // FLT_MAX, as the reference's -lowerBound entry
            std::priority_queue<P, std::vector<P>, ByPair> top;
            std::priority_queue<P, std::vector<P>, ByPairNeg> cand;
            E0++;
            if (track) seen.push_back(curdist);
            const bool ep_ok = ok(cur);
            if (ep_ok) {
                top.emplace(curdist, cur);
                curr++;
            }
            cand.emplace(-std::numeric_limits<float>::infinity(), cur);
            visited[cur] = tag;
            bool first = true;
            while (!cand.empty()) {
                P cp = cand.top();
                const float cd = first ? (ep_ok ? curdist : fmax) : cp.first;
                const P key = first ? P(curdist, cur) : cp;
                first = false;
                if ((curr == max_c && top.top() < key) || (cd > eps && curr >= min_c)) break;
                cand.pop();
                const uint32_t* data = h.ll0(cp.second);
                size_t size = Hnsw::count_of(data);
                H0++;
                for (size_t j = 1; j <= size; j++) {
                    uint32_t cid = data[j];
                    if (visited[cid] == tag) continue;
                    visited[cid] = tag;
                    float d = dist(space, dim, q, h.vec(cid));
                    E0++;
                    if (track) seen.push_back(d);
                    P np(d, cid);
                    if (curr < max_c || np < top.top()) {
                        cand.push(np);
                        if (ok(cid)) {
                            top.push(np);
                            curr++;
                        }
                        while (curr > max_c) {
                            top.pop();
                            curr--;
                        }
                    }
                }
            }
            res.resize(top.size());
            for (size_t i = res.size(); i-- > 0;) {
                res[i] = top.top();
                top.pop();
            }
        }
        // filter_results (stop_condition.h:262-269)
        while (!res.empty() && res.back().first > eps) res.pop_back();
        while (res.size() > max_c) res.pop_back();
        for (auto& r : res) out.push_back({r.first, r.second});
        if (ctr) {
            ctr[0] = H0;
            ctr[1] = E0;
            ctr[2] = Hu;
            ctr[3] = Eu;
        }
        if (tie) {
            std::sort(seen.begin(), seen.end());
            *tie = std::adjacent_find(seen.begin(), seen.end()) != seen.end() ? 1 : 0;
        }
    }
};

}  // namespace

extern "C" {

const char* or_last_error(void) { return g_err.c_str(); }

float or_distance(int space, size_t dim, const float* a, const float* b) { return dist(space, dim, a, b); }

void or_distance_batch(int space, size_t dim, const float* a, const float* b, size_t n, float* out) {
    for (size_t i = 0; i < n; i++) out[i] = dist(space, dim, a + i * dim, b + i * dim);
}

double or_distance_fp64(int space, size_t dim, const float* a, const float* b) {
    double s = 0;
    if (space == OR_SPACE_L2) {
        for (size_t i = 0; i < dim; i++) {
            double t = (double)a[i] - (double)b[i];
            s += t * t;
        }
        return s;
    }
    for (size_t i = 0; i < dim; i++) s += (double)a[i] * (double)b[i];
    return 1.0 - s;
}

void* or_hnsw_load(const char* path) {
    try {
        std::ifstream in(path, std::ios::binary);
        if (!in.is_open()) throw std::runtime_error("Cannot open file");
        std::vector<char> buf((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        return parse(buf.data(), buf.size());
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void* or_hnsw_load_bytes(const void* buf, size_t len) {
    try {
        return parse((const char*)buf, len);
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void or_hnsw_free(void* h) { delete (Hnsw*)h; }
void or_hnsw_labels(void* hv, uint64_t* labels_by_id, uint8_t* deleted_by_id) {
    const Hnsw& h = *(const Hnsw*)hv;
    for (size_t i = 0; i < h.count; i++) {
        if (labels_by_id) labels_by_id[i] = h.label((uint32_t)i);
        if (deleted_by_id) deleted_by_id[i] = h.deleted((uint32_t)i) ? 1 : 0;
    }
}
size_t or_hnsw_count(void* h) { return ((Hnsw*)h)->count; }
int or_hnsw_maxlevel(void* h) { return ((Hnsw*)h)->maxlevel; }
uint32_t or_hnsw_entry(void* h) { return ((Hnsw*)h)->entry; }

int or_hnsw_search(void* hv, int space, const float* q, size_t nq, size_t k, size_t ef,
                   const int32_t* ef_per_query, const int32_t* k_per_query, size_t k_stride, int mode,
                   int nthreads, uint64_t* labels, float* dists, int64_t* counters, uint8_t* tie_flags) {
    return or_hnsw_search_allowed(hv, space, q, nq, k, ef, ef_per_query, k_per_query, k_stride, mode, nthreads,
                                  nullptr, labels, dists, counters, tie_flags);
}

static int search_rows(void* hv, int space, const float* q, size_t nq, size_t k, size_t ef,
                       const int32_t* ef_per_query, const int32_t* k_per_query, size_t k_stride, int mode,
                       int nthreads, const uint8_t* allowed, const int32_t* hop_caps, uint64_t* labels,
                       float* dists, int64_t* counters, uint8_t* tie_flags);

int or_hnsw_search_allowed(void* hv, int space, const float* q, size_t nq, size_t k, size_t ef,
                           const int32_t* ef_per_query, const int32_t* k_per_query, size_t k_stride, int mode,
                           int nthreads, const uint8_t* allowed, uint64_t* labels, float* dists, int64_t* counters,
                           uint8_t* tie_flags) {
    return search_rows(hv, space, q, nq, k, ef, ef_per_query, k_per_query, k_stride, mode, nthreads, allowed, nullptr,
                       labels, dists, counters, tie_flags);
}

int or_hnsw_search_hopcap(void* hv, int space, const float* q, size_t nq, size_t k, size_t ef,
                          const int32_t* hop_caps, size_t k_stride, int mode, int nthreads, uint64_t* labels,
                          float* dists, int64_t* counters, uint8_t* tie_flags) {
    return search_rows(hv, space, q, nq, k, ef, nullptr, nullptr, k_stride, mode, nthreads, nullptr, hop_caps, labels,
                       dists, counters, tie_flags);
}

static int search_rows(void* hv, int space, const float* q, size_t nq, size_t k, size_t ef,
                       const int32_t* ef_per_query, const int32_t* k_per_query, size_t k_stride, int mode,
                       int nthreads, const uint8_t* allowed, const int32_t* hop_caps, uint64_t* labels,
                       float* dists, int64_t* counters, uint8_t* tie_flags) {
    const Hnsw& h = *(const Hnsw*)hv;
    if (k_stride == 0) k_stride = k;
    if (nthreads <= 0) nthreads = (int)std::thread::hardware_concurrency();
    std::vector<Searcher*> ss(std::max(nthreads, 1), nullptr);
    std::atomic<int> short_rows(0);
    std::mutex err_mu;
    bool failed = false;
    parallel_for(nq, nthreads, [&](size_t r, int t) {
        if (!ss[t]) ss[t] = new Searcher(h, space);
        size_t kq = k_per_query ? (size_t)k_per_query[r] : k;
        size_t efq = ef_per_query ? (size_t)ef_per_query[r] : ef;
        std::vector<Res> out;
        try {
            const int64_t cap = hop_caps ? (int64_t)hop_caps[r] : 0;
            if (mode == OR_MODE_HEAP)
                ss[t]->run<OR_MODE_HEAP>(q + r * h.dim, kq, efq, out, counters ? counters + 4 * r : nullptr,
                                         tie_flags ? tie_flags + r : nullptr, allowed, cap);
            else
                ss[t]->run<OR_MODE_TOTAL>(q + r * h.dim, kq, efq, out, counters ? counters + 4 * r : nullptr,
                                          tie_flags ? tie_flags + r : nullptr, allowed, cap);
        } catch (std::exception& e) {
            std::lock_guard<std::mutex> g(err_mu);
            g_err = e.what();
            failed = true;
            return;
        }
        // `out` is farthest-first; the reference result heap orders (dist, label) pairs and
        // the binding writes them closest-first (bindings.cpp:649-654, hnswalg.h:1317-1321).
        std::vector<std::pair<float, uint64_t>> lab;
        for (auto& o : out) lab.emplace_back(o.d, h.label(o.id));
        std::sort(lab.begin(), lab.end());
        for (size_t i = 0; i < kq; i++) {
            labels[r * k_stride + i] = i < lab.size() ? lab[i].second : UINT64_MAX;
            dists[r * k_stride + i] = i < lab.size() ? lab[i].first : std::numeric_limits<float>::infinity();
        }
        if (lab.size() != kq) short_rows++;
    });
    for (auto* s : ss) delete s;
    return failed ? -1 : short_rows.load();
}

int or_hnsw_search_epsilon(void* hv, int space, const float* q, size_t nq, float eps, size_t min_c, size_t max_c,
                           int mode, int nthreads, const uint8_t* allowed, size_t out_stride, uint64_t* labels,
                           float* dists, uint32_t* counts, int64_t* counters, uint8_t* tie_flags) {
    const Hnsw& h = *(const Hnsw*)hv;
    if (nthreads <= 0) nthreads = (int)std::thread::hardware_concurrency();
    std::vector<Searcher*> ss(std::max(nthreads, 1), nullptr);
    std::mutex err_mu;
    bool failed = false;
    parallel_for(nq, nthreads, [&](size_t r, int t) {
        if (!ss[t]) ss[t] = new Searcher(h, space);
        std::vector<Res> out;
        try {
            if (mode == OR_MODE_HEAP)
                ss[t]->run_eps<OR_MODE_HEAP>(q + r * h.dim, eps, min_c, max_c, out,
                                             counters ? counters + 4 * r : nullptr,
                                             tie_flags ? tie_flags + r : nullptr, allowed);
            else
                ss[t]->run_eps<OR_MODE_TOTAL>(q + r * h.dim, eps, min_c, max_c, out,
                                              counters ? counters + 4 * r : nullptr,
                                              tie_flags ? tie_flags + r : nullptr, allowed);
        } catch (std::exception& e) {
            std::lock_guard<std::mutex> g(err_mu);
            g_err = e.what();
            failed = true;
            return;
        }
        // the reference returns the result vector in heap order (closest first); rows here are
        // ordered by (distance, label) so that equal distances have one order
        std::vector<std::pair<float, uint64_t>> lab;
        for (auto& o : out) lab.emplace_back(o.d, h.label(o.id));
        if (mode == OR_MODE_TOTAL) std::sort(lab.begin(), lab.end());
        counts[r] = (uint32_t)lab.size();
        for (size_t i = 0; i < out_stride; i++) {
            labels[r * out_stride + i] = i < lab.size() ? lab[i].second : UINT64_MAX;
            dists[r * out_stride + i] = i < lab.size() ? lab[i].first : std::numeric_limits<float>::infinity();
        }
    });
    for (auto* s : ss) delete s;
    return failed ? -1 : 0;
}

int or_hnsw_search_multivector(void* hv, int space, const float* q, size_t nq, const uint32_t* docs,
                               size_t num_docs, size_t ef_collection, int mode, int nthreads, const uint8_t* allowed,
                               size_t out_stride, uint64_t* labels, float* dists, uint32_t* counts,
                               int64_t* counters, uint8_t* tie_flags) {
    const Hnsw& h = *(const Hnsw*)hv;
    if (nthreads <= 0) nthreads = (int)std::thread::hardware_concurrency();
    std::vector<Searcher*> ss(std::max(nthreads, 1), nullptr);
    std::mutex err_mu;
    bool failed = false;
    parallel_for(nq, nthreads, [&](size_t r, int t) {
        if (!ss[t]) ss[t] = new Searcher(h, space);
        std::vector<Res> out;
        try {
            if (mode == OR_MODE_HEAP)
                ss[t]->run_mv<OR_MODE_HEAP>(q + r * h.dim, docs, num_docs, ef_collection, out,
                                            counters ? counters + 4 * r : nullptr,
                                            tie_flags ? tie_flags + r : nullptr, allowed);
            else
                ss[t]->run_mv<OR_MODE_TOTAL>(q + r * h.dim, docs, num_docs, ef_collection, out,
                                             counters ? counters + 4 * r : nullptr,
                                             tie_flags ? tie_flags + r : nullptr, allowed);
        } catch (std::exception& e) {
            std::lock_guard<std::mutex> g(err_mu);
            g_err = e.what();
            failed = true;
            return;
        }
        std::vector<std::pair<float, uint64_t>> lab;
        for (auto& o : out) lab.emplace_back(o.d, h.label(o.id));
        if (mode == OR_MODE_TOTAL) std::sort(lab.begin(), lab.end());
        counts[r] = (uint32_t)lab.size();
        for (size_t i = 0; i < out_stride; i++) {
            labels[r * out_stride + i] = i < lab.size() ? lab[i].second : UINT64_MAX;
            dists[r * out_stride + i] = i < lab.size() ? lab[i].first : std::numeric_limits<float>::infinity();
        }
    });
    for (auto* s : ss) delete s;
    return failed ? -1 : 0;
}

int or_bf_search(int space, size_t dim, const float* data, const uint64_t* labels_in, size_t n,
                 const float* q, size_t nq, size_t k, int nthreads, uint64_t* labels, float* dists) {
    parallel_for(nq, nthreads, [&](size_t r, int) {
        // bruteforce.h:106-125: admit if dist <= lastdist or not full; pop the max (dist,label).
        std::priority_queue<std::pair<float, uint64_t>> top;
        float last = std::numeric_limits<float>::max();
        const float* qq = q + r * dim;
        for (size_t i = 0; i < n; i++) {
            float d = dist(space, dim, qq, data + i * dim);
            if (d <= last || top.size() < k) {
                uint64_t lab = labels_in ? labels_in[i] : (uint64_t)i;
                top.emplace(d, lab);
                if (top.size() > k) top.pop();
                if (!top.empty()) last = top.top().first;
            }
        }
        for (int i = (int)k - 1; i >= 0; i--) {
            if (top.empty()) {
                labels[r * k + i] = UINT64_MAX;
                dists[r * k + i] = std::numeric_limits<float>::infinity();
                continue;
            }
            labels[r * k + i] = top.top().second;
            dists[r * k + i] = top.top().first;
            top.pop();
        }
    });
    return 0;
}

void or_fp64_pairs(int space, size_t dim, const float* q, const float* data, const int64_t* qi,
                   const int64_t* ri, size_t n, double* out) {
    for (size_t i = 0; i < n; i++) out[i] = or_distance_fp64(space, dim, q + qi[i] * dim, data + ri[i] * dim);
}

// ---------------------------------------------------------------------------------------------
// Learned-stop MLP reference.
// ---------------------------------------------------------------------------------------------
int or_mlp_param_count(int in_dim, int hidden) { return hidden * in_dim + hidden + hidden * hidden + hidden + hidden + 1; }

static inline double gelu(double x) {
    const double c = 0.7978845608028654;  // sqrt(2/pi)
    return 0.5 * x * (1.0 + std::tanh(c * (x + 0.044715 * x * x * x)));
}
static inline double gelu_grad(double x) {
    const double c = 0.7978845608028654;
    double u = c * (x + 0.044715 * x * x * x);
    double t = std::tanh(u);
    double du = c * (1.0 + 3.0 * 0.044715 * x * x);
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * du;
}

struct MlpView {
    const float *W1, *b1, *W2, *b2, *w3, *b3;
    MlpView(const float* p, int in, int hid) {
        W1 = p;
        b1 = W1 + hid * in;
        W2 = b1 + hid;
        b2 = W2 + hid * hid;
        w3 = b2 + hid;
        b3 = w3 + hid;
    }
};

void or_mlp_forward(const float* params, int in_dim, int hidden, const float* x, size_t n, float* logits) {
    MlpView m(params, in_dim, hidden);
    std::vector<double> a1(hidden), h1(hidden), a2(hidden), h2(hidden);
    for (size_t r = 0; r < n; r++) {
        const float* xr = x + r * in_dim;
        for (int j = 0; j < hidden; j++) {
            double s = m.b1[j];
            for (int i = 0; i < in_dim; i++) s += (double)m.W1[j * in_dim + i] * xr[i];
            a1[j] = s;
            h1[j] = gelu(s);
        }
        for (int j = 0; j < hidden; j++) {
            double s = m.b2[j];
            for (int i = 0; i < hidden; i++) s += (double)m.W2[j * hidden + i] * h1[i];
            a2[j] = s;
            h2[j] = gelu(s);
        }
        double o = m.b3[0];
        for (int j = 0; j < hidden; j++) o += (double)m.w3[j] * h2[j];
        logits[r] = (float)o;
    }
}

double or_mlp_loss_grad(const float* params, int in_dim, int hidden, const float* x, const float* y, size_t n,
                        double* grad) {
    MlpView m(params, in_dim, hidden);
    int P = or_mlp_param_count(in_dim, hidden);
    std::fill(grad, grad + P, 0.0);
    double* gW1 = grad;
    double* gb1 = gW1 + hidden * in_dim;
    double* gW2 = gb1 + hidden;
    double* gb2 = gW2 + hidden * hidden;
    double* gw3 = gb2 + hidden;
    double* gb3 = gw3 + hidden;
    std::vector<double> a1(hidden), h1(hidden), a2(hidden), h2(hidden), d2(hidden), d1(hidden);
    double loss = 0;
    for (size_t r = 0; r < n; r++) {
        const float* xr = x + r * in_dim;
        for (int j = 0; j < hidden; j++) {
            double s = m.b1[j];
            for (int i = 0; i < in_dim; i++) s += (double)m.W1[j * in_dim + i] * xr[i];
            a1[j] = s;
            h1[j] = gelu(s);
        }
        for (int j = 0; j < hidden; j++) {
            double s = m.b2[j];
            for (int i = 0; i < hidden; i++) s += (double)m.W2[j * hidden + i] * h1[i];
            a2[j] = s;
            h2[j] = gelu(s);
        }
        double o = m.b3[0];
        for (int j = 0; j < hidden; j++) o += (double)m.w3[j] * h2[j];
        double yy = y[r];
        // BCE with logits: max(o,0) - o*y + log(1+exp(-|o|))
        loss += std::max(o, 0.0) - o * yy + std::log1p(std::exp(-std::fabs(o)));
        double go = (1.0 / (1.0 + std::exp(-o)) - yy) / (double)n;
        gb3[0] += go;
        for (int j = 0; j < hidden; j++) {
            gw3[j] += go * h2[j];
            d2[j] = go * m.w3[j] * gelu_grad(a2[j]);
        }
        for (int j = 0; j < hidden; j++) {
            gb2[j] += d2[j];
            for (int i = 0; i < hidden; i++) gW2[j * hidden + i] += d2[j] * h1[i];
        }
        for (int i = 0; i < hidden; i++) {
            double s = 0;
            for (int j = 0; j < hidden; j++) s += d2[j] * m.W2[j * hidden + i];
            d1[i] = s * gelu_grad(a1[i]);
        }
        for (int j = 0; j < hidden; j++) {
            gb1[j] += d1[j];
            for (int i = 0; i < in_dim; i++) gW1[j * in_dim + i] += d1[j] * xr[i];
        }
    }
    return loss / (double)n;
}

}  // extern "C"
