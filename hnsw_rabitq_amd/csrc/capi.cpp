// C ABI of the engine (declarations and reference mapping in include/hnswgpu.h).
#include "../../include/hnswgpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "host/hnsw_index.hpp"
#include "kernels.h"

#include "capi_common.h"

using namespace hgapi;

namespace {

// ------------------------------------------------------------------------------------------
// GPU copy of an HNSW graph.
// ------------------------------------------------------------------------------------------
struct GpuIndex;
// Live exported GPU indexes and their export serials (a captured graph checks that its index
// still exists before a replay).
std::mutex g_live_mu;
std::unordered_map<const GpuIndex*, uint64_t> g_live;

struct GpuIndex {
    uint64_t serial = 0;  // unique per export (filters name the export they were built for)
    int device = 0;
    int cus = 0;
    hg::DevGraph g{};
    DevBuf vec, links0, upper_off, upper, labels;
    // 8-bit coded copy of the vectors (DevGraph::vec8), made at export when it is lossless;
    // g.vec8 points at it unless the caller chose the FP32 rows (hg_gpu_set_vector_code)
    DevBuf vec8;
    bool has_u8 = false;
    // Per-launch device words (status, work counters, overflow-list counts; layout at kWords)
    // and ovf = two overflow lists of [nq] u32.  Launches take the next of kRing sets, so up to
    // kRing launches on different streams may be in flight on one index.
    // A set's event is recorded after the launch that used it; the next launch that takes the
    // set waits for it on its own stream (so a work counter is never cleared under a running
    // launch), and hg_gpu_search_status waits for it before reading the words.
    struct LaunchWords {
        DevBuf words, ovf;
        hipEvent_t done = nullptr;
        bool pending = false;
        ~LaunchWords() {
            if (done) (void)hipEventDestroy(done);
        }
    };
    static constexpr int kRing = 16;
    LaunchWords ring[kRing];
    unsigned next_ring = 0, last_ring = 0;
    // changes whenever the delete marks are re-read (a captured search graph checks it)
    uint64_t del_version = 0;
    // non-bare-bone search: delete marks as a bitmap over internal ids, and the host labels
    // (filters are built from label lists)
    DevBuf delbits;
    size_t num_deleted = 0;
    std::vector<uint64_t> host_labels;
    // host-path staging
    DevBuf q, ol, od, ctr, cnt;
    hipStream_t stream = nullptr;
    size_t bytes = 0;
    std::mutex mu;
    // tuning: beam capacity -> (hash_bits, blocks per CU)
    std::map<uint32_t, std::pair<uint32_t, int>> tuned;
    size_t urows = 0;  // rows of the upper-level link array
    ~GpuIndex() {
        {
            std::lock_guard<std::mutex> lk(g_live_mu);
            g_live.erase(this);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct Plan {
    uint32_t cap, hash_bits;
    int blocks;
    size_t lds;
};

uint32_t ceil_log2(uint64_t x) {
    uint32_t b = 0;
    while ((1ull << b) < x) b++;
    return b;
}

// Membership filter of one exported index: bit i set = the label of internal id i is allowed
// (the isIdAllowed functor of hnswalg.h:325-326, 406-407, evaluated once per element).
struct GpuFilter {
    GpuIndex* gi = nullptr;
    uint64_t serial = 0;
    int device = 0;
    size_t n = 0;
    DevBuf bits;
    size_t allowed = 0;
};

std::atomic<uint64_t> g_export_serial{0};

// Largest beam of the non-bare-bone search: every admitted key stays in the beam until the
// result set holds ef allowed entries, so a selective filter needs about ef / (allowed fraction)
// entries.  8192 entries = 64 KiB of LDS.
constexpr uint32_t kFiltMaxCap = 8192;
// Launch words: [0] status, [1..12] work counters of the passes, [13..24] overflow-list counts.
constexpr int kMaxPasses = 12, kWordStatus = 0, kWordWork = 1, kWordCount = 1 + kMaxPasses;
constexpr uint32_t kWords = 1 + 2 * kMaxPasses;

Plan make_plan(GpuIndex* gi, size_t max_ef, size_t max_k, size_t nq, bool filtered = false,
               double allowed_frac = 1.0, uint32_t min_cap = 0) {
    size_t ef_eff = std::max(max_ef, max_k);
    Plan p;
    p.cap = (uint32_t)((ef_eff + 63) / 64 * 64);
    auto it = filtered ? gi->tuned.end() : gi->tuned.find(p.cap);
    if (filtered) {
        // ef allowed entries are found after about ef / (allowed fraction) admissions; the
        // margin covers the spread of that count for elements allowed independently of their
        // position (a query whose neighbourhood holds fewer allowed elements is flagged, and the
        // host path runs it again at the largest beam)
        const double want = 1.1 * (double)ef_eff / std::max(allowed_frac, 1e-4) + 128.0;
        const double c = std::min<double>(kFiltMaxCap, std::max<double>({want, (double)p.cap, (double)min_cap}));
        p.cap = std::min<uint32_t>(kFiltMaxCap, ((uint32_t)c + 63) / 64 * 64);
        // lossy, like the tuned bare-bone plans: LDS goes to the beam
        p.hash_bits = 10;
    } else if (it != gi->tuned.end()) {
        p.hash_bits = it->second.first;
    } else {
        // the hash is lossy (an overwritten slot only costs a re-evaluation, in the same distance
        // round), so it is sized for occupancy: about one slot per four beam entries
        p.hash_bits = std::min<uint32_t>(10, std::max<uint32_t>(8, ceil_log2((ef_eff + 3) / 4)));
    }
    if (filtered) {
        // a CU holds 160 KiB of LDS: shrink the (lossy) hash, then the beam, to fit
        while (hg::search_lds_bytes(gi->g.dim, p.cap, p.hash_bits) > 160 * 1024 && p.hash_bits > 9)
            p.hash_bits--;
        while (hg::search_lds_bytes(gi->g.dim, p.cap, p.hash_bits) > 160 * 1024 && p.cap > 64) p.cap -= 64;
        require(p.cap >= ef_eff, "ef and dimension too large for the filtered GPU search");
        // the largest beam that keeps the same number of workgroups per CU (a free margin)
        const size_t per = (size_t)(160 * 1024) / hg::search_lds_bytes(gi->g.dim, p.cap, p.hash_bits);
        if (per >= 1) {
            const size_t budget = (size_t)(160 * 1024) / per - 512;
            while (p.cap + 64 <= kFiltMaxCap && hg::search_lds_bytes(gi->g.dim, p.cap + 64, p.hash_bits) <= budget)
                p.cap += 64;
        }
    }
    hg::SearchParams sp{};
    sp.g = gi->g;
    sp.cap = p.cap;
    sp.hash_bits = p.hash_bits;
    p.lds = hg::search_lds_bytes(gi->g.dim, p.cap, p.hash_bits);
    int per_cu = hg::search_blocks_per_cu(sp, filtered);
    if (it != gi->tuned.end() && it->second.second > 0) per_cu = std::min(per_cu, it->second.second);
    size_t blocks = (size_t)gi->cus * (size_t)per_cu;
    p.blocks = (int)std::max<size_t>(1, std::min(blocks, nq));
    return p;
}

// Device words of one search launch: kWords status / counter words and, for the non-bare-bone
// search, two overflow lists of nq u32 (may be null for the bare-bone search).
struct WordsView {
    uint32_t* words;
    uint32_t* ovf;
};
WordsView view(GpuIndex::LaunchWords* lw) { return WordsView{(uint32_t*)lw->words.p, (uint32_t*)lw->ovf.p}; }

hg::SearchParams base_params(GpuIndex* gi, const Plan& pl, uint32_t* words) {
    hg::SearchParams sp{};
    sp.g = gi->g;
    sp.cap = pl.cap;
    sp.hash_bits = pl.hash_bits;
    sp.work = words + kWordWork;
    sp.trace_period = 1;
    return sp;
}

// The next set of launch words of the ring (host state under the index mutex), for a launch
// on stream s: the launch waits on the device for the set's previous launch to end.
GpuIndex::LaunchWords* next_words(GpuIndex* gi, hipStream_t s) {
    gi->last_ring = gi->next_ring++ % GpuIndex::kRing;
    GpuIndex::LaunchWords* lw = &gi->ring[gi->last_ring];
    if (lw->pending) {
        hip_check(hipStreamWaitEvent(s, lw->done, 0), "hipStreamWaitEvent");
        lw->pending = false;
    }
    return lw;
}

// Marks the end of the launch that used lw on stream s.
void words_done(GpuIndex::LaunchWords* lw, hipStream_t s) {
    if (!lw->done) hip_check(hipEventCreateWithFlags(&lw->done, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventRecord(lw->done, s), "hipEventRecord");
    lw->pending = true;
}

// Fraction of the elements a query may return (not deleted, allowed by the filter), used to
// size the beam of the non-bare-bone search.
double allowed_fraction(const GpuIndex* gi, const GpuFilter* f) {
    const double n = std::max<double>(1.0, (double)gi->g.n);
    double frac = ((double)gi->g.n - (double)gi->num_deleted) / n;
    if (f) frac = std::min(frac, (double)f->allowed / n);
    return frac;
}

bool is_filtered(const GpuIndex* gi, const GpuFilter* filt) { return gi->num_deleted > 0 || filt != nullptr; }

// Range search (EpsilonSearchStopCondition, stop_condition.h:207-275).
struct EpsArgs {
    float eps;
    size_t min_c, max_c;
    uint32_t* counts;  // device [nq]
};

// Reserves the launch words (and, for the non-bare-bone search, the overflow list) that a
// launch of nq queries needs; done before a graph capture, where allocation is not allowed.
// A set whose buffers grow is first waited for (its old buffers are freed).
void prepare_words(GpuIndex::LaunchWords* lw, size_t nq, bool filtered) {
    const size_t ovf = filtered ? std::max<size_t>(nq, 1) * 8 : 0;
    if (lw->done && (lw->words.bytes < kWords * 4 || lw->ovf.bytes < ovf))
        hip_check(hipEventSynchronize(lw->done), "hipEventSynchronize");
    lw->words.reserve(kWords * 4);
    if (filtered) lw->ovf.reserve(ovf);
}

// The non-bare-bone branch runs when the index holds deleted elements or a filter is given
// (hnswalg.h:1304-1312), in passes at growing beams: the first planned from the allowed
// fraction, each later one (x1.5, up to the largest) taking the queries that outgrew the previous
// one from a list kept on the device.  All passes are asynchronous and graph-capturable.  status
// (device word, may be null = the launch words' own) gets bit 0 when a query outgrew the largest
// beam too.  lw: the launch words to use (prepared).
hipError_t run_search(GpuIndex* gi, const float* dq, size_t nq, const int32_t* def, const int32_t* dk, size_t ef,
                      size_t k, size_t k_stride, uint64_t* dl, float* dd, int32_t* dc, StopModel* stop,
                      hipStream_t s, const WordsView& lw, const GpuFilter* filt = nullptr,
                      uint32_t* status = nullptr, const EpsArgs* eps = nullptr) {
    const bool filtered = is_filtered(gi, filt) || eps != nullptr;
    const double frac = filtered ? allowed_fraction(gi, filt) : 1.0;
    // the range search's beam holds at least min_c results before it may stop: the first pass
    // is planned from that, later passes grow it
    const size_t plan_ef = eps ? std::max<size_t>(eps->min_c, 1) : ef;
    const size_t plan_k = eps ? 1 : k;
    Plan pl = make_plan(gi, plan_ef, plan_k, nq, filtered, frac);
    hg::SearchParams sp = base_params(gi, pl, lw.words);
    sp.queries = dq;
    sp.nq = (uint32_t)nq;
    sp.ef_pq = def;
    sp.k_pq = dk;
    sp.ef = (uint32_t)ef;
    sp.k = (uint32_t)k;
    sp.k_stride = (uint32_t)k_stride;
    sp.out_labels = dl;
    sp.out_dists = dd;
    sp.counters = dc;
    if (stop) {
        sp.use_stop = 1;
        sp.stop = dev_stop(stop);
    }
    if (eps) {
        sp.ef = (uint32_t)std::min<size_t>(eps->max_c, 0x7FFFFFFF);
        sp.eps_mode = 1;
        sp.eps = eps->eps;
        sp.eps_min = (int)std::min<size_t>(eps->min_c, 0x7FFFFFFF);
        sp.counts = eps->counts;
    }
    if (!filtered) {
        // the bare-bone beam is clamped to its capacity: no query can outgrow it
        if (status) {
            hipError_t e = hg::launch_zero_u32(status, 1, s);
            if (e != hipSuccess) return e;
        }
        if (!(stop && stop->ext_ef > (int)std::max(ef, k))) return hg::launch_search(sp, pl.blocks, s);
        // extended learned search: the first pass lists the queries whose network is not
        // confident at the end of their search; the second answers them at ext_ef (the plain
        // fixed-ef search, with the same stop checks), reading the list's count on the device
        hipError_t e = hg::launch_zero_u32(lw.words, kWords, s);
        if (e != hipSuccess) return e;
        sp.ext_list = lw.ovf;
        sp.ext_count = lw.words + kWordCount;
        e = hg::launch_search(sp, pl.blocks, s);
        if (e != hipSuccess) return e;
        const size_t ext = (size_t)stop->ext_ef;
        Plan p2 = make_plan(gi, ext, k, nq);
        hg::SearchParams s2 = sp;
        s2.cap = p2.cap;
        s2.hash_bits = p2.hash_bits;
        s2.work = lw.words + kWordWork + 1;
        s2.ext_list = nullptr;
        s2.ext_count = nullptr;
        s2.order = lw.ovf;
        s2.nq_dev = lw.words + kWordCount;
        s2.ef_pq = nullptr;
        s2.ef = (uint32_t)ext;
        return hg::launch_search(s2, p2.blocks, s);
    }

    uint32_t* words = lw.words;
    sp.del_bits = gi->num_deleted > 0 ? (const uint32_t*)gi->delbits.p : nullptr;
    sp.allow_bits = filt ? (const uint32_t*)filt->bits.p : nullptr;
    sp.status = status ? status : words + kWordStatus;
    hipError_t e = hg::launch_zero_u32(words, kWords, s);  // status, work counters, list counts
    if (e == hipSuccess && status) e = hg::launch_zero_u32(status, 1, s);
    if (e != hipSuccess) return e;
    // passes at growing beams: each lists the queries that outgrew it for the next (the lists
    // alternate between two halves of ovf); the last, at the largest beam, sets the status
    Plan cur = pl;
    for (int pass = 0;; pass++) {
        const bool last = cur.cap >= kFiltMaxCap || pass + 1 >= kMaxPasses;
        hg::SearchParams sq = sp;
        sq.cap = cur.cap;
        sq.hash_bits = cur.hash_bits;
        sq.work = words + kWordWork + pass;
        if (pass > 0) {
            sq.order = (const uint32_t*)lw.ovf + (size_t)((pass - 1) & 1) * nq;
            sq.nq_dev = words + kWordCount + pass - 1;
        }
        if (!last) {
            sq.ovf_list = lw.ovf + (size_t)(pass & 1) * nq;
            sq.ovf_count = words + kWordCount + pass;
        }
        e = hg::launch_search(sq, cur.blocks, s);
        if (e != hipSuccess || last) return e;
        const uint32_t next = std::min<uint32_t>(kFiltMaxCap, (uint32_t)(cur.cap * 3 / 2 + 63) / 64 * 64);
        cur = make_plan(gi, plan_ef, plan_k, nq, true, frac, next);
    }
}

// The stop model asks for a second pass at a larger ef (the overflow list holds its queries).
bool extends(void* stop, size_t ef, size_t k) {
    return stop && static_cast<StopModel*>(stop)->ext_ef > (int)std::max(ef, k);
}

void check_stop_args(const GpuIndex* gi, const GpuFilter* filt, void* stop) {
    require(!(stop && is_filtered(gi, filt)),
            "the learned stop is not available with deleted elements or a filter");
}

GpuFilter* as_filter(GpuIndex* gi, void* f) {
    if (!f) return nullptr;
    auto* gf = static_cast<GpuFilter*>(f);
    require(gf->gi == gi && gf->serial == gi->serial && gf->n == gi->host_labels.size(),
            "the filter was built for another GPU index");
    return gf;
}

void check_search_args(GpuIndex* gi, size_t nq, size_t ef, size_t k, size_t k_stride) {
    require(k >= 1, "k must be at least 1");
    require(k_stride >= k, "k_stride must be >= k");
    require(std::max(ef, k) <= 4096, "max(ef, k) above 4096 is not supported by the GPU search");
    require(nq < (1ull << 31), "too many queries");
    (void)gi;
}

}  // namespace

// ============================================================================================
HG_API const char* hg_last_error(void) { return g_err.c_str(); }

// ---- CPU index ----
HG_API int hg_hnsw_create(int space, size_t dim, size_t max_elements, size_t M, size_t ef_construction,
                          size_t random_seed, void** out) {
    return guarded([&] {
        require(out != nullptr, "null output handle");
        require(space == HG_SPACE_L2 || space == HG_SPACE_IP, "Space name must be one of l2, ip, or cosine.");
        require(dim >= 1, "dim must be positive");
        require(M >= 2, "M must be at least 2");
        *out = new hnswgpu::HnswIndex(space, dim, max_elements, M, ef_construction, random_seed);
        return HG_OK;
    });
}

HG_API int hg_hnsw_load(int space, size_t dim, const char* path, size_t max_elements, void** out) {
    return guarded([&] {
        require(out && path, "null argument");
        require(space == HG_SPACE_L2 || space == HG_SPACE_IP, "Space name must be one of l2, ip, or cosine.");
        *out = new hnswgpu::HnswIndex(space, dim, std::string(path), max_elements);
        return HG_OK;
    });
}

HG_API int hg_hnsw_add_items(void* h, const float* data, size_t rows, const uint64_t* labels, int num_threads) {
    return guarded([&] {
        require(h != nullptr, "null index");
        require(rows == 0 || data != nullptr, "null data");
        static_cast<hnswgpu::HnswIndex*>(h)->add_items(data, rows, labels, num_threads);
        return HG_OK;
    });
}

// The first step of hg_hnsw_build_gpu on its own: insertion into the levels above 0 only.
HG_API int hg_hnsw_add_items_upper(void* h, const float* data, size_t rows, const uint64_t* labels,
                                   int num_threads) {
    return guarded([&] {
        require(h != nullptr, "null index");
        require(rows == 0 || data != nullptr, "null data");
        static_cast<hnswgpu::HnswIndex*>(h)->add_items_upper(data, rows, labels, num_threads);
        return HG_OK;
    });
}

HG_API int hg_hnsw_save(void* h, const char* path) {
    return guarded([&] {
        require(h && path, "null argument");
        static_cast<hnswgpu::HnswIndex*>(h)->save(std::string(path));
        return HG_OK;
    });
}

HG_API int hg_hnsw_serialize(void* h, void* buffer, size_t capacity, size_t* size) {
    return guarded([&] {
        require(h && size, "null argument");
        auto* idx = static_cast<hnswgpu::HnswIndex*>(h);
        size_t need = idx->serialized_size();
        *size = need;
        if (buffer && capacity >= need) idx->serialize((char*)buffer);
        else if (buffer) throw Error(HG_ERR_INVALID, "buffer too small");
        return HG_OK;
    });
}

HG_API int hg_hnsw_info(void* h, int64_t* info) {
    return guarded([&] {
        require(h && info, "null argument");
        auto* idx = static_cast<hnswgpu::HnswIndex*>(h);
        info[0] = (int64_t)idx->count();
        info[1] = (int64_t)idx->max_elements();
        info[2] = (int64_t)idx->dim();
        info[3] = (int64_t)idx->M();
        info[4] = (int64_t)idx->ef_construction();
        info[5] = (int64_t)idx->maxlevel();
        info[6] = (int64_t)idx->entry();
        info[7] = (int64_t)idx->num_deleted();
        return HG_OK;
    });
}

HG_API int hg_hnsw_set_ef(void* h, size_t ef) {
    return guarded([&] {
        require(h != nullptr, "null index");
        static_cast<hnswgpu::HnswIndex*>(h)->set_ef(ef);
        return HG_OK;
    });
}

HG_API int hg_hnsw_mark_deleted(void* h, uint64_t label) {
    return guarded([&] {
        require(h != nullptr, "null index");
        static_cast<hnswgpu::HnswIndex*>(h)->mark_deleted(label);
        return HG_OK;
    });
}

HG_API int hg_hnsw_unmark_deleted(void* h, uint64_t label) {
    return guarded([&] {
        require(h != nullptr, "null index");
        static_cast<hnswgpu::HnswIndex*>(h)->unmark_deleted(label);
        return HG_OK;
    });
}

HG_API int hg_hnsw_get_ef(void* h, size_t* ef) {
    return guarded([&] {
        require(h && ef, "null argument");
        *ef = static_cast<hnswgpu::HnswIndex*>(h)->ef();
        return HG_OK;
    });
}

HG_API int hg_hnsw_free(void* h) {
    delete static_cast<hnswgpu::HnswIndex*>(h);
    return HG_OK;
}

HG_API int hg_normalize_rows(const float* in, float* out, size_t rows, size_t dim) {
    return guarded([&] {
        require((in && out) || rows == 0, "null argument");
        for (size_t r = 0; r < rows; r++) hnswgpu::normalize_vector(in + r * dim, out + r * dim, dim);
        return HG_OK;
    });
}

// ---- GPU index ----
namespace {
// The GPU layout in host memory: what an export builds from the CPU index, what the layout cache
// file holds, and what finish_export uploads.
struct HostLayout {
    size_t n = 0, dim = 0, stride = 0, maxM = 0, maxM0 = 0, urows = 0;
    int maxlevel = 0, space = 0;
    uint32_t entry = 0;
    std::vector<float> vec;
    std::vector<uint32_t> l0, uoff, up, del;
    std::vector<uint64_t> labels;
};

void check_layout(const HostLayout& L) {
    require(L.n > 0, "the index is empty");
    require(L.n < (1ull << 30), "more than 2^30 elements are not supported by the GPU search");
    // one wave holds a whole neighbour list: 2M <= 64 (M <= 32)
    require(L.maxM0 <= 64 && L.maxM <= 64, "M above 32 is not supported by the GPU search");
    require(L.dim >= 1 && L.dim <= 4096, "dim above 4096 is not supported by the GPU search");
    require(L.stride == (L.dim + 3) / 4 * 4 && L.entry < L.n, "Index seems to be corrupted or unsupported");
}

GpuIndex* finish_export(HostLayout&& L, int device) {
    check_layout(L);
    int ndev = 0;
    hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    require(device >= 0 && device < ndev, "invalid device");
    DeviceGuard dg(device);
    std::unique_ptr<GpuIndex> gi(new GpuIndex());
    gi->device = device;
    gi->serial = ++g_export_serial;
    // one attribute (hipGetDeviceProperties fills every field and costs a large fraction of a second)
    hip_check(hipDeviceGetAttribute(&gi->cus, hipDeviceAttributeMultiprocessorCount, device),
              "hipDeviceGetAttribute");
    auto upload = [&](DevBuf& b, const void* src, size_t bytes) {
        b.reserve(std::max<size_t>(bytes, 16));
        hip_check(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
        gi->bytes += bytes;
    };
    upload(gi->vec, L.vec.data(), L.vec.size() * 4);
    upload(gi->links0, L.l0.data(), L.l0.size() * 4);
    upload(gi->upper_off, L.uoff.data(), L.uoff.size() * 4);
    upload(gi->upper, L.up.data(), L.up.size() * 4);
    upload(gi->labels, L.labels.data(), L.labels.size() * 8);
    upload(gi->delbits, L.del.data(), L.del.size() * 4);
    for (uint32_t w : L.del) gi->num_deleted += (size_t)__builtin_popcount(w);
    if (hg::search_fixed_chunks((uint32_t)L.dim)) {
        // the 8-bit copy, kept only when every coordinate is an integer in 0..255
        DevBuf flag;
        flag.reserve(4);
        gi->vec8.reserve(std::max<size_t>(L.n * L.dim, 16));
        hip_check(hipMemset(flag.p, 0, 4), "hipMemset");
        hip_check(hg::launch_code_u8((const float*)gi->vec.p, (uint32_t)L.n, (uint32_t)L.dim, (uint32_t)L.stride,
                                     (uint8_t*)gi->vec8.p, (uint32_t*)flag.p, nullptr),
                  "code launch");
        uint32_t bad = 1;
        hip_check(hipMemcpy(&bad, flag.p, 4, hipMemcpyDeviceToHost), "hipMemcpy D2H");
        if (bad == 0) {
            gi->has_u8 = true;
            gi->bytes += L.n * L.dim;
        } else {
            hip_check(hipFree(gi->vec8.p), "hipFree");
            gi->vec8.p = nullptr;
            gi->vec8.bytes = 0;
        }
    }
    for (auto& lw : gi->ring) lw.words.reserve(kWords * 4);
    gi->host_labels = std::move(L.labels);
    gi->urows = L.urows;
    hip_check(hipStreamCreateWithFlags(&gi->stream, hipStreamNonBlocking), "hipStreamCreate");
    hg::DevGraph& g = gi->g;
    g.vec = (const float*)gi->vec.p;
    g.vec8 = gi->has_u8 ? (const uint8_t*)gi->vec8.p : nullptr;
    g.links0 = (const uint32_t*)gi->links0.p;
    g.upper_off = (const uint32_t*)gi->upper_off.p;
    g.upper = (const uint32_t*)gi->upper.p;
    g.labels = (const uint64_t*)gi->labels.p;
    g.n = (uint32_t)L.n;
    g.dim = (uint32_t)L.dim;
    g.stride = (uint32_t)L.stride;
    g.maxM = (uint32_t)L.maxM;
    g.maxM0 = (uint32_t)L.maxM0;
    g.maxlevel = L.maxlevel;
    g.entry = L.entry;
    g.space = L.space;
    {
        std::lock_guard<std::mutex> lk(g_live_mu);
        g_live[gi.get()] = gi->serial;
    }
    return gi.release();
}

// Layout cache file: magic, version, the layout's sizes, its arrays, then an FNV-1a 64-bit hash
// of everything before it.
constexpr uint32_t kLayoutMagic = 0x4C474748u;  // "HGGL"
constexpr uint32_t kLayoutVersion = 1;

struct Fnv {
    uint64_t h = 1469598103934665603ull;
    void add(const void* p, size_t n) {
        const unsigned char* c = (const unsigned char*)p;
        for (size_t i = 0; i < n; i++) h = (h ^ c[i]) * 1099511628211ull;
    }
};
}  // namespace

HG_API int hg_gpu_index_create(void* hv, int device, void** out) {
    return guarded([&] {
        require(hv && out, "null argument");
        auto* h = static_cast<hnswgpu::HnswIndex*>(hv);
        HostLayout L;
        L.n = h->count();
        L.dim = h->dim();
        L.stride = (L.dim + 3) / 4 * 4;
        L.maxM = h->maxM();
        L.maxM0 = h->maxM0();
        L.maxlevel = h->maxlevel();
        L.entry = h->entry();
        L.space = h->space();
        check_layout(L);
        const size_t n = L.n, dim = L.dim, stride = L.stride, maxM = L.maxM, maxM0 = L.maxM0;
        // Export (integrity checked as in checkIntegrity, hnswalg.h:1380-1409: ids in range).
        L.vec.assign(n * stride, 0.f);
        L.l0.assign(n * maxM0, hg::SENT);
        L.uoff.assign(n, hg::SENT);
        L.labels.resize(n);
        L.del.assign((n + 31) / 32, 0u);
        size_t urows = 0;
        for (size_t i = 0; i < n; i++) {
            if (h->element_level((uint32_t)i) > 0) {
                L.uoff[i] = (uint32_t)urows;
                urows += (size_t)h->element_level((uint32_t)i);
            }
        }
        L.urows = urows;
        L.up.assign(std::max<size_t>(urows, 1) * maxM, hg::SENT);
        for (size_t i = 0; i < n; i++) {
            memcpy(&L.vec[i * stride], h->vector((uint32_t)i), dim * 4);
            L.labels[i] = h->label((uint32_t)i);
            if (h->is_deleted((uint32_t)i)) L.del[i >> 5] |= 1u << (i & 31);
            const uint32_t* ll = h->linklist0((uint32_t)i);
            size_t c = hnswgpu::HnswIndex::list_count(ll);
            if (c > maxM0) throw Error(HG_ERR_RUNTIME, "a link list holds more entries than its capacity");
            for (size_t j = 0; j < c; j++) {
                if (ll[1 + j] >= n) throw Error(HG_ERR_RUNTIME, "cand error");
                L.l0[i * maxM0 + j] = ll[1 + j];
            }
            for (int lv = 1; lv <= h->element_level((uint32_t)i); lv++) {
                const uint32_t* lu = h->linklist((uint32_t)i, lv);
                size_t cu = hnswgpu::HnswIndex::list_count(lu);
                if (cu > maxM) throw Error(HG_ERR_RUNTIME, "a link list holds more entries than its capacity");
                for (size_t j = 0; j < cu; j++) {
                    if (lu[1 + j] >= n) throw Error(HG_ERR_RUNTIME, "cand error");
                    if (h->element_level(lu[1 + j]) < lv)
                        throw Error(HG_ERR_RUNTIME, "a link on a level above the element's own");
                    L.up[(L.uoff[i] + (size_t)(lv - 1)) * maxM + j] = lu[1 + j];
                }
            }
        }
        require(h->element_level(h->entry()) >= h->maxlevel(), "entry point level mismatch");
        *out = finish_export(std::move(L), device);
        return HG_OK;
    });
}

// Saves the exported layout (the device arrays as they are, delete marks included).
HG_API int hg_gpu_index_save(void* gv, const char* path) {
    return guarded([&] {
        require(gv && path, "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        std::lock_guard<std::mutex> lk(gi->mu);
        DeviceGuard dg(gi->device);
        const hg::DevGraph& g = gi->g;
        const size_t n = g.n;
        HostLayout L;
        L.vec.resize(n * g.stride);
        L.l0.resize(n * g.maxM0);
        L.uoff.resize(n);
        L.up.resize(std::max<size_t>(gi->urows, 1) * g.maxM);
        L.del.resize((n + 31) / 32);
        hip_check(hipMemcpy(L.vec.data(), gi->vec.p, L.vec.size() * 4, hipMemcpyDeviceToHost), "D2H");
        hip_check(hipMemcpy(L.l0.data(), gi->links0.p, L.l0.size() * 4, hipMemcpyDeviceToHost), "D2H");
        hip_check(hipMemcpy(L.uoff.data(), gi->upper_off.p, L.uoff.size() * 4, hipMemcpyDeviceToHost), "D2H");
        hip_check(hipMemcpy(L.up.data(), gi->upper.p, L.up.size() * 4, hipMemcpyDeviceToHost), "D2H");
        hip_check(hipMemcpy(L.del.data(), gi->delbits.p, L.del.size() * 4, hipMemcpyDeviceToHost), "D2H");
        std::ofstream f(path, std::ios::binary);
        require((bool)f, std::string("Cannot open file ") + path);
        Fnv fnv;
        auto put = [&](const void* p, size_t bytes) {
            f.write((const char*)p, (std::streamsize)bytes);
            fnv.add(p, bytes);
        };
        const uint32_t magic = kLayoutMagic, version = kLayoutVersion;
        const uint64_t hdr[7] = {n, g.dim, g.stride, g.maxM, g.maxM0, gi->urows, 0};
        const int32_t hdr2[3] = {g.maxlevel, (int32_t)g.entry, g.space};
        put(&magic, 4);
        put(&version, 4);
        put(hdr, sizeof(hdr));
        put(hdr2, sizeof(hdr2));
        put(L.vec.data(), L.vec.size() * 4);
        put(L.l0.data(), L.l0.size() * 4);
        put(L.uoff.data(), L.uoff.size() * 4);
        put(L.up.data(), L.up.size() * 4);
        put(gi->host_labels.data(), n * 8);
        put(L.del.data(), L.del.size() * 4);
        f.write((const char*)&fnv.h, 8);
        require((bool)f, "write failed");
        return HG_OK;
    });
}

// Loads a saved layout straight to the device (no CPU index).
HG_API int hg_gpu_index_load(const char* path, int device, void** out) {
    return guarded([&] {
        require(path && out, "null argument");
        std::ifstream f(path, std::ios::binary);
        require((bool)f, std::string("Cannot open file ") + path);
        Fnv fnv;
        auto get = [&](void* p, size_t bytes) {
            f.read((char*)p, (std::streamsize)bytes);
            require((bool)f, "Index seems to be corrupted or unsupported");
            fnv.add(p, bytes);
        };
        uint32_t magic = 0, version = 0;
        get(&magic, 4);
        get(&version, 4);
        require(magic == kLayoutMagic, "not a GPU layout file");
        require(version == kLayoutVersion, "GPU layout file of another version: export the index again");
        uint64_t hdr[7];
        int32_t hdr2[3];
        get(hdr, sizeof(hdr));
        get(hdr2, sizeof(hdr2));
        HostLayout L;
        L.n = hdr[0];
        L.dim = hdr[1];
        L.stride = hdr[2];
        L.maxM = hdr[3];
        L.maxM0 = hdr[4];
        L.urows = hdr[5];
        L.maxlevel = hdr2[0];
        L.entry = (uint32_t)hdr2[1];
        L.space = hdr2[2];
        check_layout(L);
        require(L.urows <= L.n * 64 && (L.space == 0 || L.space == 1), "Index seems to be corrupted or unsupported");
        {
            // the arrays the header announces must be the file's remaining bytes (no allocation
            // sized by a damaged header)
            const std::streamoff here = f.tellg();
            f.seekg(0, f.end);
            const size_t left = (size_t)(f.tellg() - here);
            f.seekg(here, f.beg);
            const size_t need = (L.n * L.stride + L.n * L.maxM0 + L.n + std::max<size_t>(L.urows, 1) * L.maxM) * 4 +
                                L.n * 8 + (L.n + 31) / 32 * 4 + 8;
            require(left == need, "Index seems to be corrupted or unsupported");
        }
        L.vec.resize(L.n * L.stride);
        L.l0.resize(L.n * L.maxM0);
        L.uoff.resize(L.n);
        L.up.resize(std::max<size_t>(L.urows, 1) * L.maxM);
        L.labels.resize(L.n);
        L.del.resize((L.n + 31) / 32);
        get(L.vec.data(), L.vec.size() * 4);
        get(L.l0.data(), L.l0.size() * 4);
        get(L.uoff.data(), L.uoff.size() * 4);
        get(L.up.data(), L.up.size() * 4);
        get(L.labels.data(), L.labels.size() * 8);
        get(L.del.data(), L.del.size() * 4);
        uint64_t stored = 0;
        f.read((char*)&stored, 8);
        require((bool)f && stored == fnv.h, "Index seems to be corrupted or unsupported");
        // the kernels index with these unchecked: ids in range, upper rows only of nodes that
        // have some, and maxlevel rows of padding so that no level can read past the array
        for (uint32_t v : L.l0) require(v == hg::SENT || v < L.n, "cand error");
        for (uint32_t v : L.uoff) require(v == hg::SENT || v < L.urows, "Index seems to be corrupted or unsupported");
        for (uint32_t v : L.up) require(v == hg::SENT || (v < L.n && L.uoff[v] != hg::SENT), "cand error");
        require(L.maxlevel >= 0 && L.maxlevel < 64, "Index seems to be corrupted or unsupported");
        require(L.maxlevel == 0 || L.uoff[L.entry] != hg::SENT, "entry point level mismatch");
        L.up.resize((std::max<size_t>(L.urows, 1) + (size_t)L.maxlevel) * L.maxM, hg::SENT);
        *out = finish_export(std::move(L), device);
        return HG_OK;
    });
}

HG_API int hg_gpu_index_free(void* g) {
    if (!g) return HG_OK;
    auto* gi = static_cast<GpuIndex*>(g);
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(gi->device);
    delete gi;
    if (prev >= 0) (void)hipSetDevice(prev);
    return HG_OK;
}

HG_API int hg_gpu_set_vector_code(void* gv, int code) {
    return guarded([&] {
        require(gv, "null argument");
        require(code == HG_VEC_AUTO || code == HG_VEC_F32, "code must be HG_VEC_AUTO or HG_VEC_F32");
        auto* gi = static_cast<GpuIndex*>(gv);
        std::lock_guard<std::mutex> lk(gi->mu);
        const uint8_t* want = (code == HG_VEC_AUTO && gi->has_u8) ? (const uint8_t*)gi->vec8.p : nullptr;
        if (want != gi->g.vec8) {
            gi->g.vec8 = want;
            gi->tuned.clear();  // plans were timed with the other kernel
        }
        return HG_OK;
    });
}

HG_API int hg_gpu_vector_code(void* gv, int* code) {
    return guarded([&] {
        require(gv && code, "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        std::lock_guard<std::mutex> lk(gi->mu);
        *code = gi->g.vec8 ? HG_VEC_U8 : HG_VEC_F32;
        return HG_OK;
    });
}

HG_API int hg_gpu_index_bytes(void* g, size_t* bytes) {
    return guarded([&] {
        require(g && bytes, "null argument");
        *bytes = static_cast<GpuIndex*>(g)->bytes;
        return HG_OK;
    });
}

namespace {
// Host-buffer search (the knn_query path): upload, launch, download.  If a query of the
// non-bare-bone search outgrows even the largest beam, the call fails rather than return a list
// that may differ from the reference's.
int search_host(GpuIndex* gi, const float* queries, size_t nq, size_t k, size_t ef, const GpuFilter* filt,
                uint64_t* labels, float* dists, int32_t* counters) {
    check_search_args(gi, nq, ef, k, k);
    if (nq == 0) return HG_OK;
    require(queries != nullptr, "null queries");
    DeviceGuard dg(gi->device);
    const size_t dim = gi->g.dim;
    gi->q.reserve(nq * dim * 4);
    gi->ol.reserve(nq * k * 8);
    gi->od.reserve(nq * k * 4);
    if (counters) gi->ctr.reserve(nq * 16);
    hipStream_t s = gi->stream;
    const bool filtered = gi->num_deleted > 0 || filt != nullptr;
    hip_check(hipMemcpyAsync(gi->q.p, queries, nq * dim * 4, hipMemcpyHostToDevice, s), "hipMemcpy H2D");
    GpuIndex::LaunchWords* lw = next_words(gi, s);
    prepare_words(lw, nq, filtered);
    hip_check(run_search(gi, (const float*)gi->q.p, nq, nullptr, nullptr, ef, k, k, (uint64_t*)gi->ol.p,
                         (float*)gi->od.p, counters ? (int32_t*)gi->ctr.p : nullptr, nullptr, s, view(lw), filt),
              "search launch");
    words_done(lw, s);
    if (filtered) {
        uint32_t status = 0;
        hip_check(hipMemcpyAsync(&status, (uint32_t*)lw->words.p + kWordStatus, 4, hipMemcpyDeviceToHost, s),
                  "hipMemcpy D2H");
        hip_check(hipStreamSynchronize(s), "search");
        if (status & 1u) {
            g_err = "filtered search: a query needed more than " + std::to_string(kFiltMaxCap) +
                    " beam entries (the allowed elements are too few for this ef)";
            return HG_ERR_RUNTIME;
        }
    }
    hip_check(hipMemcpyAsync(labels, gi->ol.p, nq * k * 8, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
    hip_check(hipMemcpyAsync(dists, gi->od.p, nq * k * 4, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
    if (counters) hip_check(hipMemcpyAsync(counters, gi->ctr.p, nq * 16, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
    hip_check(hipStreamSynchronize(s), "search");
    for (size_t r = 0; r < nq; r++)
        if (labels[r * k + k - 1] == ~0ull) {
            g_err = "Cannot return the results in a contiguous 2D array. Probably ef or M is too small";
            return HG_ERR_SHORT;
        }
    return HG_OK;
}
}  // namespace

HG_API int hg_gpu_search(void* gv, const float* queries, size_t nq, size_t k, size_t ef, uint64_t* labels,
                         float* dists, int32_t* counters) {
    return guarded([&] {
        require(gv && labels && dists, "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        std::lock_guard<std::mutex> lk(gi->mu);
        return search_host(gi, queries, nq, k, ef, nullptr, labels, dists, counters);
    });
}

HG_API int hg_gpu_search_filtered(void* gv, const float* queries, size_t nq, size_t k, size_t ef, void* filter,
                                  uint64_t* labels, float* dists, int32_t* counters) {
    return guarded([&] {
        require(gv && labels && dists, "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        std::lock_guard<std::mutex> lk(gi->mu);
        return search_host(gi, queries, nq, k, ef, as_filter(gi, filter), labels, dists, counters);
    });
}

namespace {
void check_eps_args(float eps, size_t min_c, size_t max_c, size_t out_stride) {
    require(max_c >= 1 && min_c <= max_c, "need 1 <= max_candidates and min_candidates <= max_candidates");
    require(out_stride >= 1 && out_stride <= kFiltMaxCap, "out_stride must be in [1, 8192]");
    require(!(eps != eps), "epsilon is NaN");
}
}  // namespace

// searchStopConditionClosest with EpsilonSearchStopCondition on host buffers (hnswalg.h:1326-1377).
HG_API int hg_gpu_search_epsilon(void* gv, const float* queries, size_t nq, float epsilon, size_t min_candidates,
                                 size_t max_candidates, void* filter, size_t out_stride, uint64_t* labels,
                                 float* dists, uint32_t* counts) {
    return guarded([&] {
        require(gv && labels && dists && counts, "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        check_eps_args(epsilon, min_candidates, max_candidates, out_stride);
        GpuFilter* gf = as_filter(gi, filter);
        require(nq < (1ull << 31), "too many queries");
        if (nq == 0) return HG_OK;
        require(queries != nullptr, "null queries");
        std::lock_guard<std::mutex> lk(gi->mu);
        DeviceGuard dg(gi->device);
        const size_t dim = gi->g.dim;
        gi->q.reserve(nq * dim * 4);
        gi->ol.reserve(nq * out_stride * 8);
        gi->od.reserve(nq * out_stride * 4);
        gi->cnt.reserve(nq * 4);
        hipStream_t s = gi->stream;
        hip_check(hipMemcpyAsync(gi->q.p, queries, nq * dim * 4, hipMemcpyHostToDevice, s), "hipMemcpy H2D");
        GpuIndex::LaunchWords* lw = next_words(gi, s);
        prepare_words(lw, nq, true);
        EpsArgs ea{epsilon, min_candidates, max_candidates, (uint32_t*)gi->cnt.p};
        hip_check(run_search(gi, (const float*)gi->q.p, nq, nullptr, nullptr, max_candidates, out_stride, out_stride,
                             (uint64_t*)gi->ol.p, (float*)gi->od.p, nullptr, nullptr, s, view(lw), gf, nullptr, &ea),
                  "search launch");
        words_done(lw, s);
        uint32_t status = 0;
        hip_check(hipMemcpyAsync(&status, (uint32_t*)lw->words.p + kWordStatus, 4, hipMemcpyDeviceToHost, s),
                  "hipMemcpy D2H");
        hip_check(hipMemcpyAsync(labels, gi->ol.p, nq * out_stride * 8, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
        hip_check(hipMemcpyAsync(dists, gi->od.p, nq * out_stride * 4, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
        hip_check(hipMemcpyAsync(counts, gi->cnt.p, nq * 4, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
        hip_check(hipStreamSynchronize(s), "search");
        if (status & 1u) {
            g_err = "range search: a query needed more than " + std::to_string(kFiltMaxCap) +
                    " beam entries (lower min_candidates or epsilon)";
            return HG_ERR_RUNTIME;
        }
        return HG_OK;
    });
}

// The same on device buffers, asynchronous on stream (d_status as hg_gpu_search_device_filtered).
HG_API int hg_gpu_search_epsilon_device(void* gv, const float* dq, size_t nq, float epsilon, size_t min_candidates,
                                        size_t max_candidates, void* filter, size_t out_stride, uint64_t* dl,
                                        float* dd, uint32_t* d_counts, int32_t* dc, uint32_t* d_status,
                                        void* stream) {
    return guarded([&] {
        require(gv && dl && dd && d_counts, "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        check_eps_args(epsilon, min_candidates, max_candidates, out_stride);
        GpuFilter* gf = as_filter(gi, filter);
        require(nq < (1ull << 31), "too many queries");
        if (nq == 0) return HG_OK;
        require(dq != nullptr, "null queries");
        std::lock_guard<std::mutex> lk(gi->mu);
        DeviceGuard dg(gi->device);
        GpuIndex::LaunchWords* lw = next_words(gi, (hipStream_t)stream);
        prepare_words(lw, nq, true);
        EpsArgs ea{epsilon, min_candidates, max_candidates, d_counts};
        hip_check(run_search(gi, dq, nq, nullptr, nullptr, max_candidates, out_stride, out_stride, dl, dd, dc, nullptr,
                             (hipStream_t)stream, view(lw), gf, d_status, &ea),
                  "search launch");
        words_done(lw, (hipStream_t)stream);
        return HG_OK;
    });
}

HG_API int hg_gpu_filter_create(void* gv, const uint64_t* allowed_labels, size_t n_allowed, void** out) {
    return guarded([&] {
        require(gv && out && (n_allowed == 0 || allowed_labels), "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        std::vector<uint64_t> sorted(allowed_labels, allowed_labels + n_allowed);
        std::sort(sorted.begin(), sorted.end());
        const size_t n = gi->host_labels.size();
        std::vector<uint32_t> bits((n + 31) / 32, 0u);
        std::unique_ptr<GpuFilter> f(new GpuFilter());
        for (size_t i = 0; i < n; i++)
            if (std::binary_search(sorted.begin(), sorted.end(), gi->host_labels[i])) {
                bits[i >> 5] |= 1u << (i & 31);
                f->allowed++;
            }
        DeviceGuard dg(gi->device);
        f->gi = gi;
        f->serial = gi->serial;
        f->device = gi->device;
        f->n = n;
        f->bits.reserve(std::max<size_t>(bits.size() * 4, 16));
        hip_check(hipMemcpy(f->bits.p, bits.data(), bits.size() * 4, hipMemcpyHostToDevice), "hipMemcpy H2D");
        *out = f.release();
        return HG_OK;
    });
}

HG_API int hg_gpu_filter_free(void* f) {
    if (!f) return HG_OK;
    auto* gf = static_cast<GpuFilter*>(f);
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(gf->device);
    delete gf;
    if (prev >= 0) (void)hipSetDevice(prev);
    return HG_OK;
}

HG_API int hg_gpu_filter_count(void* f, size_t* allowed) {
    return guarded([&] {
        require(f && allowed, "null argument");
        *allowed = static_cast<GpuFilter*>(f)->allowed;
        return HG_OK;
    });
}

HG_API int hg_gpu_search_status(void* gv, uint32_t* status, uint32_t* second_pass_queries) {
    return guarded([&] {
        require(gv != nullptr, "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        std::lock_guard<std::mutex> lk(gi->mu);
        DeviceGuard dg(gi->device);
        uint32_t w[kWords] = {};
        GpuIndex::LaunchWords& lw = gi->ring[gi->last_ring];
        if (!lw.words.p) return HG_OK;  // no search yet
        // the words of the last launch, read once that launch has ended
        if (lw.pending) hip_check(hipEventSynchronize(lw.done), "hipEventSynchronize");
        hip_check(hipMemcpy(w, lw.words.p, sizeof(w), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        if (status) *status = w[kWordStatus];
        if (second_pass_queries) *second_pass_queries = w[kWordCount];
        return HG_OK;
    });
}

// Re-reads the delete marks of the CPU index the GPU index was exported from.
HG_API int hg_gpu_refresh_deleted(void* gv, void* cpu) {
    return guarded([&] {
        require(gv && cpu, "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        auto* h = static_cast<hnswgpu::HnswIndex*>(cpu);
        const size_t n = gi->host_labels.size();
        require(h->count() == n, "the CPU index changed size since the export; export it again");
        std::vector<uint32_t> del((n + 31) / 32, 0u);
        size_t nd = 0;
        for (size_t i = 0; i < n; i++) {
            require(h->label((uint32_t)i) == gi->host_labels[i], "the CPU index changed since the export");
            if (h->is_deleted((uint32_t)i)) {
                del[i >> 5] |= 1u << (i & 31);
                nd++;
            }
        }
        std::lock_guard<std::mutex> lk(gi->mu);
        DeviceGuard dg(gi->device);
        hip_check(hipMemcpy(gi->delbits.p, del.data(), del.size() * 4, hipMemcpyHostToDevice), "hipMemcpy H2D");
        gi->num_deleted = nd;
        gi->del_version++;
        return HG_OK;
    });
}

HG_API int hg_gpu_index_labels(void* gv, uint64_t* labels, size_t capacity, size_t* count) {
    return guarded([&] {
        require(gv && count, "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        *count = gi->host_labels.size();
        if (labels) {
            require(capacity >= gi->host_labels.size(), "label buffer too small");
            std::copy(gi->host_labels.begin(), gi->host_labels.end(), labels);
        }
        return HG_OK;
    });
}

HG_API int hg_gpu_search_device(void* gv, const float* dq, size_t nq, const int32_t* def, const int32_t* dk, size_t ef,
                                size_t k, size_t k_stride, uint64_t* dl, float* dd, int32_t* dc, void* stop,
                                void* stream) {
    return guarded([&] {
        require(gv && dl && dd, "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        check_search_args(gi, nq, ef, k, k_stride);
        if (nq == 0) return HG_OK;
        require(dq != nullptr, "null queries");
        check_stop_args(gi, nullptr, stop);
        DeviceGuard dg(gi->device);
        std::lock_guard<std::mutex> lk(gi->mu);
        GpuIndex::LaunchWords* lw = next_words(gi, (hipStream_t)stream);
        prepare_words(lw, nq, is_filtered(gi, nullptr) || extends(stop, ef, k));
        hip_check(run_search(gi, dq, nq, def, dk, ef, k, k_stride, dl, dd, dc, static_cast<StopModel*>(stop),
                             (hipStream_t)stream, view(lw)),
                  "search launch");
        words_done(lw, (hipStream_t)stream);
        return HG_OK;
    });
}

HG_API int hg_gpu_search_device_filtered(void* gv, const float* dq, size_t nq, const int32_t* def,
                                         const int32_t* dk, size_t ef, size_t k, size_t k_stride, uint64_t* dl,
                                         float* dd, int32_t* dc, void* filter, uint32_t* d_status, void* stream) {
    return guarded([&] {
        require(gv && dl && dd, "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        check_search_args(gi, nq, ef, k, k_stride);
        if (nq == 0) return HG_OK;
        require(dq != nullptr, "null queries");
        GpuFilter* gf = as_filter(gi, filter);
        DeviceGuard dg(gi->device);
        std::lock_guard<std::mutex> lk(gi->mu);
        GpuIndex::LaunchWords* lw = next_words(gi, (hipStream_t)stream);
        prepare_words(lw, nq, is_filtered(gi, gf));
        hip_check(run_search(gi, dq, nq, def, dk, ef, k, k_stride, dl, dd, dc, nullptr, (hipStream_t)stream, view(lw),
                             gf, d_status),
                  "search launch");
        words_done(lw, (hipStream_t)stream);
        return HG_OK;
    });
}

// The bare-bone device search with a per-query cap on level-0 expansions (d_hop_cap, device
// int32 [nq]; <= 0 = no cap): the search driven by a stop condition that ends it after a given
// number of hops (hnswalg.h:343-355, hnswlib.h:134-150).  It replays a learned search's stopping
// points without the network, so the two can be compared hop for hop.
HG_API int hg_gpu_search_device_hopcap(void* gv, const float* dq, size_t nq, const int32_t* d_hop_cap, size_t ef,
                                       size_t k, uint64_t* dl, float* dd, int32_t* dc, void* stream) {
    return guarded([&] {
        require(gv && dl && dd && d_hop_cap, "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        check_search_args(gi, nq, ef, k, k);
        require(!is_filtered(gi, nullptr), "the hop-capped search is the bare-bone search: the index has deleted elements");
        if (nq == 0) return HG_OK;
        require(dq != nullptr, "null queries");
        DeviceGuard dg(gi->device);
        std::lock_guard<std::mutex> lk(gi->mu);
        hipStream_t s = (hipStream_t)stream;
        GpuIndex::LaunchWords* lw = next_words(gi, s);
        prepare_words(lw, nq, false);
        Plan pl = make_plan(gi, ef, k, nq);
        hg::SearchParams sp = base_params(gi, pl, (uint32_t*)lw->words.p);
        sp.queries = dq;
        sp.nq = (uint32_t)nq;
        sp.ef = (uint32_t)ef;
        sp.k = (uint32_t)k;
        sp.k_stride = (uint32_t)k;
        sp.out_labels = dl;
        sp.out_dists = dd;
        sp.counters = dc;
        sp.hop_cap = d_hop_cap;
        hip_check(hg::launch_search(sp, pl.blocks, s), "search launch");
        words_done(lw, s);
        return HG_OK;
    });
}

// ---- allocation-free device search ----
// Scratch of hg_gpu_search_device_noalloc: 64 launch words (status, work counters, overflow-list
// counts) and the two overflow lists of the non-bare-bone search (2 x kNoAllocMaxQueries u32):
// 2 MiB + 256 B in a file-scope device array, which every device the library runs on holds
// until the process ends.  Host code never uses the array's name as an address: the address of
// each device's copy comes from hipGetSymbolAddress, once per device.
constexpr size_t kNoAllocMaxQueries = 262144;
constexpr size_t kNoAllocWordSlots = 64;
static_assert(kWords <= kNoAllocWordSlots, "launch words fit their slots");
static __device__ __align__(256) uint32_t scratch[kNoAllocWordSlots + 2 * kNoAllocMaxQueries];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

// The device address of this device's `scratch` (the device must be current).
uint32_t* scratch_on(int device) {
    require(device >= 0 && device < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        uint32_t* ptr = nullptr;
        hip_check(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}
}  // namespace

HG_API int hg_gpu_search_device_noalloc(void* gv, const float* dq, size_t nq, const int32_t* def, const int32_t* dk,
                                        size_t ef, size_t k, size_t k_stride, uint64_t* dl, float* dd, int32_t* dc,
                                        void* filter, uint32_t* d_status, void* stream) {
    return guarded([&] {
        require(gv && dl && dd, "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        check_search_args(gi, nq, ef, k, k_stride);
        if (nq == 0) return HG_OK;
        require(dq != nullptr, "null queries");
        GpuFilter* gf = as_filter(gi, filter);
        require(!is_filtered(gi, gf) || nq <= kNoAllocMaxQueries,
                "allocation-free search: at most 262,144 queries per call when the index holds deleted elements or "
                "a filter is given (the overflow lists live in a fixed device array)");
        DeviceGuard dg(gi->device);
        std::lock_guard<std::mutex> lk(gi->mu);
        uint32_t* sc = scratch_on(gi->device);
        hip_check(run_search(gi, dq, nq, def, dk, ef, k, k_stride, dl, dd, dc, nullptr, (hipStream_t)stream,
                             WordsView{sc, sc + kNoAllocWordSlots}, gf, d_status),
                  "search launch");
        return HG_OK;
    });
}

// Trace collection for the learned stop: the plain search, recording every `period`
// expansions the 12 traversal features and the current top-k ids.
HG_API int hg_gpu_collect_traces(void* gv, const float* dq, size_t nq, size_t ef, size_t k, int period,
                                 int max_records, float* d_features, uint32_t* d_topk, int32_t* d_nrec,
                                 uint32_t* d_final, int end_record, void* stream) {
    return guarded([&] {
        require(gv && dq && d_features && d_topk && d_nrec && d_final, "null argument");
        require(period >= 1 && max_records >= 1, "period and max_records must be positive");
        auto* gi = static_cast<GpuIndex*>(gv);
        check_search_args(gi, nq, ef, k, k);
        require(gi->num_deleted == 0, "traces of the learned stop need an index without deleted elements");
        if (nq == 0) return HG_OK;
        DeviceGuard dg(gi->device);
        std::lock_guard<std::mutex> lk(gi->mu);
        hipStream_t s = (hipStream_t)stream;
        DevBuf ol, od;
        ol.reserve(nq * k * 8);
        od.reserve(nq * k * 4);
        Plan pl = make_plan(gi, ef, k, nq);
        GpuIndex::LaunchWords* lw = next_words(gi, s);
        prepare_words(lw, nq, false);
        hg::SearchParams sp = base_params(gi, pl, (uint32_t*)lw->words.p);
        sp.queries = dq;
        sp.nq = (uint32_t)nq;
        sp.ef = (uint32_t)ef;
        sp.k = (uint32_t)k;
        sp.k_stride = (uint32_t)k;
        sp.out_labels = (uint64_t*)ol.p;
        sp.out_dists = (float*)od.p;
        sp.trace_feat = d_features;
        sp.trace_topk = d_topk;
        sp.trace_nrec = d_nrec;
        sp.trace_final = d_final;
        sp.trace_period = period;
        sp.trace_max_rec = max_records;
        sp.trace_end = end_record ? 1 : 0;
        hip_check(hg::launch_search(sp, pl.blocks, s), "trace launch");
        words_done(lw, s);
        hip_check(hipStreamSynchronize(s), "trace search");
        return HG_OK;
    });
}

HG_API int hg_gpu_search_plan(void* gv, size_t max_ef, size_t max_k, size_t nq, int32_t* out) {
    return guarded([&] {
        require(gv && out, "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        DeviceGuard dg(gi->device);
        Plan p = make_plan(gi, max_ef, max_k, nq);
        out[0] = 1;
        out[1] = (int32_t)(1u << p.hash_bits);
        out[2] = (int32_t)p.cap;
        out[3] = (int32_t)p.lds;
        out[4] = p.blocks;
        return HG_OK;
    });
}

// Autotuner: times (hash size, workgroups per CU) candidates on the caller's queries and keeps
// the fastest for this beam capacity.
HG_API int hg_gpu_autotune(void* gv, const float* dq, size_t nq, size_t ef, size_t k, void* stream) {
    return guarded([&] {
        require(gv && dq, "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        DeviceGuard dg(gi->device);
        std::lock_guard<std::mutex> lk(gi->mu);
        hipStream_t s = (hipStream_t)stream;
        require(gi->num_deleted == 0, "the autotuner times the bare-bone search: the index has deleted elements");
        DevBuf ol, od;
        ol.reserve(nq * k * 8);
        od.reserve(nq * k * 4);
        size_t ef_eff = std::max(ef, k);
        uint32_t cap = (uint32_t)((ef_eff + 63) / 64 * 64);
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        float best = 1e30f;
        std::pair<uint32_t, int> best_cfg(0, 0);
        gi->tuned.erase(cap);
        for (uint32_t bits = 7; bits <= 13; bits++) {
            int last_blocks = -1;
            for (int per_cu : {4, 6, 8, 9, 10, 12, 16, 24, 32}) {
                gi->tuned[cap] = std::make_pair(bits, per_cu);
                Plan pl = make_plan(gi, ef, k, nq);
                if (pl.blocks == last_blocks) continue;  // capped by occupancy: already timed
                last_blocks = pl.blocks;
                float ms_sum = 0;
                for (int rep = 0; rep < 2; rep++) {
                    hip_check(hipEventRecord(e0, s), "hipEventRecord");
                    GpuIndex::LaunchWords* lw = next_words(gi, s);
                    prepare_words(lw, nq, false);
                    hip_check(run_search(gi, dq, nq, nullptr, nullptr, ef, k, k, (uint64_t*)ol.p, (float*)od.p,
                                         nullptr, nullptr, s, view(lw)),
                              "search launch");
                    words_done(lw, s);
                    hip_check(hipEventRecord(e1, s), "hipEventRecord");
                    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                    float ms = 0;
                    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                    if (rep > 0) ms_sum += ms;
                }
                if (ms_sum < best) {
                    best = ms_sum;
                    best_cfg = std::make_pair(bits, per_cu);
                }
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        gi->tuned[cap] = best_cfg;
        return HG_OK;
    });
}

// Fixes the launch plan for the beam capacity of (ef, k), as the autotuner would store it.
HG_API int hg_gpu_set_plan(void* gv, size_t ef, size_t k, uint32_t hash_bits, int workgroups_per_cu) {
    return guarded([&] {
        require(gv != nullptr, "null index");
        auto* gi = static_cast<GpuIndex*>(gv);
        require(k >= 1 && ef >= 1, "ef and k must be positive");
        require(hash_bits >= 7 && hash_bits <= 16, "hash_bits must be in [7, 16]");
        require(workgroups_per_cu >= 1 && workgroups_per_cu <= 32, "workgroups_per_cu must be in [1, 32]");
        std::lock_guard<std::mutex> lk(gi->mu);
        size_t ef_eff = std::max(ef, k);
        uint32_t cap = (uint32_t)((ef_eff + 63) / 64 * 64);
        require(hg::search_lds_bytes(gi->g.dim, cap, hash_bits) <= 160 * 1024,
                "beam and visited hash exceed the 160 KiB of LDS of a CU");
        gi->tuned[cap] = std::make_pair(hash_bits, workgroups_per_cu);
        return HG_OK;
    });
}

// External labels of internal ids on the device (trace records and final top-k carry internal ids).
HG_API int hg_gpu_id_labels(void* gv, const uint32_t* d_ids, size_t n, uint64_t* d_labels, void* stream) {
    return guarded([&] {
        require(gv && (n == 0 || (d_ids && d_labels)), "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        DeviceGuard dg(gi->device);
        hip_check(hg::launch_id_labels(gi->g, d_ids, n, d_labels, (hipStream_t)stream), "label gather launch");
        return HG_OK;
    });
}

// ---- multi-vector search ----
namespace {
// Document of every element of one export, by internal id.
struct GpuDocs {
    uint64_t serial = 0;
    int device = 0;
    size_t n = 0;
    DevBuf docs;
};

GpuDocs* as_docs(GpuIndex* gi, void* d) {
    require(d != nullptr, "null documents");
    auto* gd = static_cast<GpuDocs*>(d);
    require(gd->serial == gi->serial && gd->n == gi->host_labels.size(),
            "the documents were built for another GPU index");
    return gd;
}

hipError_t run_mv(GpuIndex* gi, const float* dq, size_t nq, const GpuDocs* gd, size_t num_docs, size_t ef_coll,
                  const GpuFilter* filt, size_t out_stride, uint64_t* dl, float* dd, uint32_t* dcounts, int32_t* dc,
                  uint32_t* status, uint32_t cap, GpuIndex::LaunchWords* lw, hipStream_t s) {
    hg::MvParams mp{};
    mp.g = gi->g;
    mp.queries = dq;
    mp.nq = (uint32_t)nq;
    mp.work = (uint32_t*)lw->words.p + kWordWork;
    mp.docs = (const uint32_t*)gd->docs.p;
    mp.del_bits = gi->num_deleted > 0 ? (const uint32_t*)gi->delbits.p : nullptr;
    mp.allow_bits = filt ? (const uint32_t*)filt->bits.p : nullptr;
    mp.num_docs = (uint32_t)num_docs;
    mp.ef_coll = (uint32_t)std::max(ef_coll, num_docs);  // MultiVectorSearchStopCondition's ef_collection_
    mp.cap = cap;
    mp.hash_bits = 11;
    mp.dcap = (uint32_t)((mp.ef_coll + 2 + 63) / 64 * 64);
    mp.out_labels = dl;
    mp.out_dists = dd;
    mp.counts = dcounts;
    mp.k_stride = (uint32_t)out_stride;
    mp.counters = dc;
    mp.status = status;
    require(hg::mv_lds_bytes(gi->g.dim, mp.cap, mp.hash_bits, mp.dcap) <= 160 * 1024,
            "ef_collection and dimension too large for the multi-vector GPU search");
    const size_t blocks = std::min<size_t>((size_t)gi->cus * (size_t)hg::mv_blocks_per_cu(mp), nq);
    hipError_t e = hg::launch_zero_u32(status, 1, s);
    if (e != hipSuccess) return e;
    return hg::launch_mv_search(mp, (int)std::max<size_t>(blocks, 1), s);
}

void check_mv_args(size_t num_docs, size_t ef_coll, size_t out_stride) {
    require(num_docs >= 1, "num_docs must be at least 1");
    require(std::max(num_docs, ef_coll) <= 1024, "ef_collection above 1024 is not supported by the GPU search");
    require(out_stride >= 1 && out_stride <= kFiltMaxCap, "out_stride must be in [1, 8192]");
}
}  // namespace

HG_API int hg_gpu_docs_create(void* gv, const uint64_t* labels, const uint32_t* docs, size_t count, void** out) {
    return guarded([&] {
        require(gv && out && (count == 0 || (labels && docs)), "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        std::unordered_map<uint64_t, uint32_t> doc_of;
        doc_of.reserve(count * 2);
        for (size_t i = 0; i < count; i++) doc_of[labels[i]] = docs[i];
        const size_t n = gi->host_labels.size();
        std::vector<uint32_t> by_id(n);
        for (size_t i = 0; i < n; i++) {
            auto it = doc_of.find(gi->host_labels[i]);
            require(it != doc_of.end(), "an element of the index has no document");
            by_id[i] = it->second;
        }
        DeviceGuard dg(gi->device);
        std::unique_ptr<GpuDocs> d(new GpuDocs());
        d->serial = gi->serial;
        d->device = gi->device;
        d->n = n;
        d->docs.reserve(std::max<size_t>(n, 1) * 4);
        hip_check(hipMemcpy(d->docs.p, by_id.data(), n * 4, hipMemcpyHostToDevice), "hipMemcpy H2D");
        *out = d.release();
        return HG_OK;
    });
}

HG_API int hg_gpu_docs_free(void* d) {
    if (!d) return HG_OK;
    auto* gd = static_cast<GpuDocs*>(d);
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(gd->device);
    delete gd;
    if (prev >= 0) (void)hipSetDevice(prev);
    return HG_OK;
}

// searchStopConditionClosest with MultiVectorSearchStopCondition on host buffers.
HG_API int hg_gpu_search_multivector(void* gv, const float* queries, size_t nq, void* docs, size_t num_docs,
                                     size_t ef_collection, void* filter, size_t out_stride, uint64_t* labels,
                                     float* dists, uint32_t* counts) {
    return guarded([&] {
        require(gv && labels && dists && counts, "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        check_mv_args(num_docs, ef_collection, out_stride);
        GpuDocs* gd = as_docs(gi, docs);
        GpuFilter* gf = as_filter(gi, filter);
        require(nq < (1ull << 31), "too many queries");
        if (nq == 0) return HG_OK;
        require(queries != nullptr, "null queries");
        std::lock_guard<std::mutex> lk(gi->mu);
        DeviceGuard dg(gi->device);
        const size_t dim = gi->g.dim;
        gi->q.reserve(nq * dim * 4);
        gi->ol.reserve(nq * out_stride * 8);
        gi->od.reserve(nq * out_stride * 4);
        gi->cnt.reserve(nq * 4);
        hipStream_t s = gi->stream;
        hip_check(hipMemcpyAsync(gi->q.p, queries, nq * dim * 4, hipMemcpyHostToDevice, s), "hipMemcpy H2D");
        uint32_t status = 0;
        for (uint32_t cap : {2048u, kFiltMaxCap}) {  // the larger beam only if a query outgrew the first
            GpuIndex::LaunchWords* lw = next_words(gi, s);
            prepare_words(lw, nq, false);
            uint32_t* st = (uint32_t*)lw->words.p + kWordStatus;
            hip_check(run_mv(gi, (const float*)gi->q.p, nq, gd, num_docs, ef_collection, gf, out_stride,
                             (uint64_t*)gi->ol.p, (float*)gi->od.p, (uint32_t*)gi->cnt.p, nullptr, st, cap, lw, s),
                      "multi-vector search launch");
            words_done(lw, s);
            hip_check(hipMemcpyAsync(&status, st, 4, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
            hip_check(hipStreamSynchronize(s), "search");
            if (!(status & 1u)) break;
        }
        if (status & 1u) {
            g_err = "multi-vector search: a query needed more than " + std::to_string(kFiltMaxCap) +
                    " beam entries";
            return HG_ERR_RUNTIME;
        }
        hip_check(hipMemcpyAsync(labels, gi->ol.p, nq * out_stride * 8, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
        hip_check(hipMemcpyAsync(dists, gi->od.p, nq * out_stride * 4, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
        hip_check(hipMemcpyAsync(counts, gi->cnt.p, nq * 4, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
        hip_check(hipStreamSynchronize(s), "search");
        return HG_OK;
    });
}

// The same on device buffers, asynchronous on stream, at one beam capacity (cap = 0: 2048);
// d_status bit 0 reports a query that outgrew it.
HG_API int hg_gpu_search_multivector_device(void* gv, const float* dq, size_t nq, void* docs, size_t num_docs,
                                            size_t ef_collection, void* filter, size_t out_stride, uint64_t* dl,
                                            float* dd, uint32_t* d_counts, int32_t* dc, uint32_t* d_status,
                                            uint32_t cap, void* stream) {
    return guarded([&] {
        require(gv && dl && dd && d_counts && d_status, "null argument");
        auto* gi = static_cast<GpuIndex*>(gv);
        check_mv_args(num_docs, ef_collection, out_stride);
        GpuDocs* gd = as_docs(gi, docs);
        GpuFilter* gf = as_filter(gi, filter);
        if (cap == 0) cap = 2048;
        require(cap % 64 == 0 && cap <= kFiltMaxCap, "cap must be a multiple of 64 up to 8192");
        if (nq == 0) return HG_OK;
        require(dq != nullptr, "null queries");
        std::lock_guard<std::mutex> lk(gi->mu);
        DeviceGuard dg(gi->device);
        GpuIndex::LaunchWords* lw = next_words(gi, (hipStream_t)stream);
        prepare_words(lw, nq, false);
        hip_check(run_mv(gi, dq, nq, gd, num_docs, ef_collection, gf, out_stride, dl, dd, d_counts, dc, d_status, cap,
                         lw, (hipStream_t)stream),
                  "multi-vector search launch");
        words_done(lw, (hipStream_t)stream);
        return HG_OK;
    });
}

// ---- level-0 construction on the GPU ----
// addPoint (hnswalg.h:1152-1266) for a batch of rows: the CPU inserts them into the levels
// above 0 (those do not depend on level 0), then the GPU builds their level-0 links in batches
// that grow with the graph (build.hip), and the lists are written back into the CPU index.
// Fresh builds of at least this many rows also build the levels above 0 on the GPU.
static size_t g_gpu_upper_min_rows = 250000;

// Level-0 lists of the elements [s0, count) of h (their upper levels already linked) on the GPU.
static void gpu_build_level0(hnswgpu::HnswIndex* h, size_t s0, int device) {
    {
        const size_t n = h->count();
        void* gv = nullptr;
        int rc = hg_gpu_index_create(h, device, &gv);
        if (rc != HG_OK) throw Error(rc, hg_last_error());
        std::unique_ptr<GpuIndex, int (*)(void*)> gi(static_cast<GpuIndex*>(gv), hg_gpu_index_free);
        DeviceGuard dg(gi->device);
        hipStream_t s = gi->stream;
        const uint32_t efc = (uint32_t)h->ef_construction();
        const uint32_t M = (uint32_t)h->M();
        const uint32_t maxM0 = gi->g.maxM0;
        const size_t bmax = 16384;
        DevBuf ident, cand_ids, cand_d, sel, group, err;
        {
            std::vector<uint64_t> id(n);
            for (size_t i = 0; i < n; i++) id[i] = i;
            ident.reserve(n * 8);
            hip_check(hipMemcpy(ident.p, id.data(), n * 8, hipMemcpyHostToDevice), "hipMemcpy H2D");
        }
        cand_ids.reserve(bmax * efc * 8);
        cand_d.reserve(bmax * efc * 4);
        sel.reserve(bmax * M * 4);
        const size_t group_bytes = hg::build_group_scratch_bytes(bmax * M);
        group.reserve(group_bytes);
        err.reserve(256);
        hip_check(hg::launch_zero_u32((uint32_t*)err.p, 1, s), "zero launch");
        std::vector<int> level(n);
        for (size_t i = 0; i < n; i++) level[i] = h->element_level((uint32_t)i);
        // entry point of the graph of elements [0, s): the first element of the highest level
        uint32_t entry = 0;
        int maxl = level[0];
        size_t seen = 1;
        GpuIndex::LaunchWords* lw = next_words(gi.get(), s);
        prepare_words(lw, bmax, false);
        for (size_t sb = std::max<size_t>(s0, 1); sb < n;) {
            // at least 256 rows: a small batch costs one beam search's latency whatever its size,
            // and the rows of a batch see one another exactly (brute force in the select kernel)
            const size_t B = std::min(std::min(bmax, std::max<size_t>(256, sb / 16)), n - sb);
            for (; seen < sb; seen++)
                if (level[seen] > maxl) {
                    maxl = level[seen];
                    entry = (uint32_t)seen;
                }
            // beam search of the batch over the graph of [0, sb)
            Plan pl = make_plan(gi.get(), efc, efc, B);
            hg::SearchParams sp = base_params(gi.get(), pl, (uint32_t*)lw->words.p);
            sp.g.entry = entry;
            sp.g.maxlevel = maxl;
            sp.g.labels = (const uint64_t*)ident.p;
            sp.queries = gi->g.vec + sb * gi->g.stride;
            sp.q_stride = gi->g.stride;
            sp.nq = (uint32_t)B;
            sp.ef = efc;
            sp.k = efc;
            sp.k_stride = efc;
            sp.out_labels = (uint64_t*)cand_ids.p;
            sp.out_dists = (float*)cand_d.p;
            sp.id_limit = (uint32_t)sb;
            hip_check(hg::launch_search(sp, pl.blocks, s), "build search launch");
            hg::BuildArgs ba{};
            ba.g = gi->g;
            ba.links0 = (uint32_t*)gi->links0.p;
            ba.s = (uint32_t)sb;
            ba.B = (uint32_t)B;
            ba.efc = efc;
            ba.M = M;
            ba.cand_ids = (const uint64_t*)cand_ids.p;
            ba.cand_d = (const float*)cand_d.p;
            ba.sel = (uint32_t*)sel.p;
            hip_check(hg::launch_build(ba, false, s), "build select launch");
            // reverse links, grouped on the device by the element that receives them (a stable
            // sort: new elements in ascending order within a receiver); no host round trip
            hip_check(hg::build_group_links(ba, group.p, group_bytes, (uint32_t*)err.p, s), "build link grouping");
            hip_check(hg::launch_build(ba, true, s), "build link launch");
            sb += B;
        }
        uint32_t bad = 0;
        hip_check(hipMemcpyAsync(&bad, err.p, 4, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
        hip_check(hipStreamSynchronize(s), "build");
        require(bad == 0, "cand error");
        // level-0 lists back into the CPU index
        std::vector<uint32_t> l0(n * maxM0);
        hip_check(hipMemcpyAsync(l0.data(), gi->links0.p, l0.size() * 4, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
        hip_check(hipStreamSynchronize(s), "build");
        for (size_t i = 0; i < n; i++) {
            const uint32_t* r = &l0[i * maxM0];
            size_t c = 0;
            while (c < maxM0 && r[c] != hg::SENT) c++;
            h->set_links0((uint32_t)i, r, c);
        }
    }
}

// Rows into h with every level built on the GPU where it pays.  The elements of level >= 1 form
// an HNSW of their own one level down, whose level 0 follows the reference's upper-level rules
// when its lists hold maxM = M links: it is built by this same function (recursively, down to
// a size the CPU inserts quickly), its lists are copied back as levels >= 1, and then level 0 is
// built over the whole index.  levels: optional given levels of the rows.
static void build_levels_gpu(hnswgpu::HnswIndex* h, const float* data, size_t rows, const uint64_t* labels,
                             const int* levels, int num_threads, int device) {
    const size_t s0 = h->count();
    if (s0 != 0 || rows < g_gpu_upper_min_rows) {
        h->add_items_upper(data, rows, labels, num_threads, levels);
        gpu_build_level0(h, s0, device);
        return;
    }
    std::vector<int> lv = levels ? std::vector<int>(levels, levels + rows) : h->draw_levels(rows);
    h->register_items(data, rows, labels, lv.data(), num_threads);
    const size_t n = h->count();
    std::vector<uint32_t> up;  // elements of level >= 1, in id order
    int top = 0;
    for (size_t i = 0; i < n; i++) {
        const int l = h->element_level((uint32_t)i);
        if (l >= 1) up.push_back((uint32_t)i);
        top = std::max(top, l);
    }
    if (up.empty()) {
        h->set_entry(0, 0);
    } else {
        const size_t m = up.size(), dim = h->dim();
        hnswgpu::HnswIndex sub(h->space(), dim, m, h->M(), h->ef_construction(), 100, h->maxM());
        std::vector<float> sv(m * dim);
        std::vector<int> sl(m);
        for (size_t k = 0; k < m; k++) {
            memcpy(&sv[k * dim], h->vector(up[k]), dim * sizeof(float));
            sl[k] = h->element_level(up[k]) - 1;
        }
        build_levels_gpu(&sub, sv.data(), m, nullptr, sl.data(), num_threads, device);
        // element ids of the sub-index are its insertion order (labels 0..m-1 in row order, ids
        // assigned as the rows register); map them through the labels
        std::vector<uint32_t> pos(m);
        for (size_t k = 0; k < m; k++) pos[k] = up[(size_t)sub.label((uint32_t)k)];
        std::vector<uint32_t> ids(std::max(h->maxM(), h->maxM0()));
        for (size_t k = 0; k < m; k++) {
            const uint32_t e = pos[k];
            const int le = h->element_level(e);
            for (int l = 1; l <= le; l++) {
                const uint32_t* src = l == 1 ? sub.linklist0((uint32_t)k) : sub.linklist((uint32_t)k, l - 1);
                const size_t c = hnswgpu::HnswIndex::list_count(src);
                for (size_t j = 0; j < c; j++) ids[j] = pos[src[1 + j]];
                h->set_links(e, l, ids.data(), c);
            }
        }
        h->set_entry(pos[sub.entry()], top);
    }
    gpu_build_level0(h, 0, device);
}

HG_API int hg_hnsw_build_gpu_upper_min_rows(size_t rows) {
    g_gpu_upper_min_rows = rows;
    return HG_OK;
}

HG_API int hg_hnsw_build_gpu(void* hv, const float* data, size_t rows, const uint64_t* labels, int num_threads,
                             int device) {
    return guarded([&] {
        require(hv != nullptr, "null index");
        require(rows == 0 || data != nullptr, "null data");
        auto* h = static_cast<hnswgpu::HnswIndex*>(hv);
        require(h->num_deleted() == 0, "the GPU build needs an index without deleted elements");
        require(h->ef_construction() >= 1 && h->ef_construction() <= 1024,
                "the GPU build supports ef_construction up to 1024");
        require(h->maxM0() <= 64, "M above 32 is not supported by the GPU search");
        if (rows == 0) return HG_OK;
        build_levels_gpu(h, data, rows, labels, nullptr, num_threads, device);
        return HG_OK;
    });
}

// ---- hipGraph of a device search ----
// A captured graph replays the kernels chosen at capture (bare-bone or non-bare-bone, from the
// delete marks then) over that export's buffers.  It records which index, export and delete-mark
// version it was captured from; a replay after the index was freed or its delete marks re-read
// fails instead of replaying a stale choice.
struct GraphHandle {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    GpuIndex::LaunchWords lw;  // the graph's own launch words (replays reuse them)
    GpuIndex* gi = nullptr;
    uint64_t serial = 0, del_version = 0;
    ~GraphHandle() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

HG_API int hg_graph_capture_search(void* gv, const float* dq, size_t nq, const int32_t* def, const int32_t* dk,
                                   size_t ef, size_t k, size_t k_stride, uint64_t* dl, float* dd, int32_t* dc,
                                   void* stop, void* stream, void** out_graph) {
    return guarded([&] {
        require(gv && dq && dl && dd && out_graph, "null argument");
        require(stream != nullptr, "graph capture needs a non-default stream");
        auto* gi = static_cast<GpuIndex*>(gv);
        check_search_args(gi, nq, ef, k, k_stride);
        require(nq > 0, "empty batch");
        DeviceGuard dg(gi->device);
        std::lock_guard<std::mutex> lk(gi->mu);
        hipStream_t s = (hipStream_t)stream;
        check_stop_args(gi, nullptr, stop);
        std::unique_ptr<GraphHandle> gh(new GraphHandle());
        // allocate and resolve kernel attributes outside the capture
        const bool filtered = is_filtered(gi, nullptr);
        prepare_words(&gh->lw, nq, filtered || extends(stop, ef, k));
        Plan pl = make_plan(gi, ef, k, nq, filtered, allowed_fraction(gi, nullptr));
        (void)pl;
        hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        hipError_t le = hipSuccess;
        try {
            le = run_search(gi, dq, nq, def, dk, ef, k, k_stride, dl, dd, dc, static_cast<StopModel*>(stop), s,
                            view(&gh->lw));
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hipGraph_t graph = nullptr;
        hipError_t ce = hipStreamEndCapture(s, &graph);
        hip_check(le, "search launch (capture)");
        hip_check(ce, "hipStreamEndCapture");
        gh->graph = graph;
        hip_check(hipGraphInstantiate(&gh->exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        gh->gi = gi;
        gh->serial = gi->serial;
        gh->del_version = gi->del_version;
        *out_graph = gh.release();
        return HG_OK;
    });
}

HG_API int hg_graph_launch(void* ghv, void* stream) {
    return guarded([&] {
        require(ghv != nullptr, "null graph");
        auto* gh = static_cast<GraphHandle*>(ghv);
        {
            std::lock_guard<std::mutex> lk(g_live_mu);
            auto it = g_live.find(gh->gi);
            if (it == g_live.end() || it->second != gh->serial)
                throw Error(HG_ERR_RUNTIME, "the GPU index of this graph was freed; capture the search again");
        }
        if (gh->gi->del_version != gh->del_version)
            throw Error(HG_ERR_RUNTIME,
                        "the delete marks changed since this graph was captured; capture the search again");
        hip_check(hipGraphLaunch(gh->exec, (hipStream_t)stream), "hipGraphLaunch");
        return HG_OK;
    });
}

HG_API int hg_graph_free(void* gh) {
    delete static_cast<GraphHandle*>(gh);
    return HG_OK;
}

// ---- device plumbing ----
HG_API int hg_device_count(int* count) {
    return guarded([&] {
        require(count != nullptr, "null argument");
        hip_check(hipGetDeviceCount(count), "hipGetDeviceCount");
        return HG_OK;
    });
}
HG_API int hg_set_device(int device) {
    return guarded([&] {
        hip_check(hipSetDevice(device), "hipSetDevice");
        return HG_OK;
    });
}
HG_API int hg_malloc(void** ptr, size_t bytes) {
    return guarded([&] {
        require(ptr != nullptr, "null argument");
        hip_check(hipMalloc(ptr, std::max<size_t>(bytes, 1)), "hipMalloc");
        return HG_OK;
    });
}
HG_API int hg_free(void* ptr) {
    return guarded([&] {
        if (ptr) hip_check(hipFree(ptr), "hipFree");
        return HG_OK;
    });
}
HG_API int hg_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
    return guarded([&] {
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream), "hipMemcpy H2D");
        hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
        return HG_OK;
    });
}
HG_API int hg_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
    return guarded([&] {
        // results may still be produced by work on another (non-blocking) stream
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream), "hipMemcpy D2H");
        hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
        return HG_OK;
    });
}
HG_API int hg_memset(void* dst, int value, size_t bytes, void* stream) {
    return guarded([&] {
        hip_check(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream), "hipMemsetAsync");
        return HG_OK;
    });
}
HG_API int hg_copy_kernel(void* dst, const void* src, size_t bytes, void* stream) {
    return guarded([&] {
        require(dst && src && bytes % 16 == 0, "copy needs device pointers and a multiple of 16 bytes");
        hip_check(hg::launch_copy(dst, src, bytes, (hipStream_t)stream), "copy launch");
        return HG_OK;
    });
}
HG_API int hg_stream_create(void** stream) {
    return guarded([&] {
        require(stream != nullptr, "null argument");
        hipStream_t s;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        *stream = s;
        return HG_OK;
    });
}
HG_API int hg_stream_destroy(void* stream) {
    return guarded([&] {
        if (stream) hip_check(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
        return HG_OK;
    });
}
HG_API int hg_stream_synchronize(void* stream) {
    return guarded([&] {
        hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
        return HG_OK;
    });
}
HG_API int hg_device_synchronize(void) {
    return guarded([&] {
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        return HG_OK;
    });
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};
HG_API int hg_timer_create(void** t) {
    return guarded([&] {
        require(t != nullptr, "null argument");
        std::unique_ptr<Timer> tm(new Timer());
        hip_check(hipEventCreate(&tm->a), "hipEventCreate");
        hip_check(hipEventCreate(&tm->b), "hipEventCreate");
        *t = tm.release();
        return HG_OK;
    });
}
HG_API int hg_timer_destroy(void* t) {
    delete static_cast<Timer*>(t);
    return HG_OK;
}
HG_API int hg_timer_start(void* t, void* stream) {
    return guarded([&] {
        require(t != nullptr, "null timer");
        hip_check(hipEventRecord(static_cast<Timer*>(t)->a, (hipStream_t)stream), "hipEventRecord");
        return HG_OK;
    });
}
HG_API int hg_timer_stop(void* t, void* stream, float* ms) {
    return guarded([&] {
        require(t && ms, "null argument");
        Timer* tm = static_cast<Timer*>(t);
        hip_check(hipEventRecord(tm->b, (hipStream_t)stream), "hipEventRecord");
        hip_check(hipEventSynchronize(tm->b), "hipEventSynchronize");
        hip_check(hipEventElapsedTime(ms, tm->a, tm->b), "hipEventElapsedTime");
        return HG_OK;
    });
}
