// C ABI: brute-force index (BruteforceSearch / BFIndex on the GPU).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/hnswgpu.h"
#include "capi_common.h"
#include "kernels.h"

using namespace hgapi;

namespace {

struct BfIndex {
    int device = 0;
    int space = 0;
    size_t dim = 0, max_elements = 0, count = 0;
    DevBuf vec, labels, vec_lp, scale, scratch, cand;
    DevBuf q, ol, od;
    DevBuf xstat;  // quantisation statistics of the rows' reduced-precision copy (3 floats)
    // labels ascending with their rows (the FP64 report maps labels to rows), for row version
    // sorted_version
    DevBuf sorted_labels, sorted_rows;
    uint64_t sorted_version = ~0ull;
    hg::LowpReport last;  // what the last reduced-precision search did besides the re-rank
    int lp_kind = 0;  // precision of vec_lp (0 = not built)
    int tuned[3] = {-1, -1, -1};  // GEMM variant chosen by the autotuner, per precision
    size_t lp_count = 0;
    hipStream_t stream = nullptr;
    std::unordered_map<uint64_t, size_t> label_row;
    std::vector<uint64_t> host_labels;  // label of each row (filters are bitmaps over rows)
    uint64_t serial = 0;   // unique per index
    uint64_t version = 0;  // changes whenever rows move or are added
    ~BfIndex() {
        if (stream) (void)hipStreamDestroy(stream);
    }
    hg::BfData data() const {
        hg::BfData d{};
        d.vec = (const float*)vec.p;
        d.labels = (const uint64_t*)labels.p;
        d.vec_bf16 = lp_kind == HG_PREC_BF16 ? (const uint16_t*)vec_lp.p : nullptr;
        d.vec_fp8 = lp_kind == HG_PREC_FP8 ? (const uint8_t*)vec_lp.p : nullptr;
        d.fp8_scale = (const float*)scale.p;
        d.norms = nullptr;
        d.allow = nullptr;
        d.n = (uint32_t)count;
        d.dim = (uint32_t)dim;
        d.space = space;
        return d;
    }
};

std::atomic<uint64_t> g_bf_serial{0};

// Label filter of one brute-force index: bitmap over its rows, valid for the row version it
// was built from (adds and deletions move rows).
struct BfFilter {
    uint64_t serial = 0, version = 0;
    int device = 0;
    size_t allowed = 0;
    DevBuf bits;
};

BfFilter* as_bf_filter(BfIndex* b, void* f) {
    if (!f) return nullptr;
    auto* bf = static_cast<BfFilter*>(f);
    require(bf->serial == b->serial, "the filter was built for another brute-force index");
    require(bf->version == b->version, "the index changed since the filter was built; build it again");
    return bf;
}

// Exact FP32 search with device buffers on stream s.
void exact_search(BfIndex* b, const float* dq, size_t nq, size_t k, uint64_t* dl, float* dd, hipStream_t s,
                  const uint32_t* allow = nullptr) {
    hg::BfData d = b->data();
    d.allow = allow;
    size_t sb = hg::bf_exact_scratch_bytes((uint32_t)nq, (uint32_t)k, d.n);
    b->scratch.reserve(sb);
    hip_check(hg::launch_bf_exact(d, dq, (uint32_t)nq, (uint32_t)k, dl, dd, b->scratch.p, s), "brute-force launch");
}

// Build (or rebuild after adds) the reduced-precision copy of the rows.
void ensure_lowp(BfIndex* b, int precision, hipStream_t s) {
    if (b->lp_kind == precision && b->lp_count == b->count) return;
    size_t elem = precision == HG_PREC_BF16 ? 2 : 1;
    b->vec_lp.reserve(std::max<size_t>(b->max_elements, 1) * b->dim * elem);
    b->scale.reserve(std::max<size_t>(b->max_elements, 1) * 4);
    hg::BfData d = b->data();
    hip_check(hg::launch_bf_convert(d, precision == HG_PREC_BF16 ? (uint16_t*)b->vec_lp.p : nullptr,
                                    precision == HG_PREC_FP8 ? (uint8_t*)b->vec_lp.p : nullptr, (float*)b->scale.p,
                                    nullptr, s),
              "convert launch");
    b->lp_kind = precision;
    b->lp_count = b->count;
    // the maxima the certification bound needs (row norms and quantisation error)
    b->xstat.reserve(64);
    hip_check(hg::launch_zero_u32((uint32_t*)b->xstat.p, 4, s), "zero launch");
    hip_check(hg::launch_quant_stats(b->data(), precision, (uint32_t*)b->xstat.p, s), "quantisation statistics");
}

// Reduced-precision search with certification and the FP64 near-tie check; the report is kept
// for hg_bf_lowp_report.
void lowp_search(BfIndex* b, const float* dq, size_t nq, size_t k, int precision, size_t rerank_c, uint64_t* dl,
                 float* dd, double* stats, hipStream_t s, int variant) {
    require(b->dim % 64 == 0, "the reduced-precision GEMM needs dim to be a multiple of 64");
    require(b->dim <= 8192, "the reduced-precision path supports dim up to 8192");
    size_t c = std::max(rerank_c, k);
    require(c <= 4096, "rerank_c above 4096 is not supported");
    c = std::min(c, b->count);
    ensure_lowp(b, precision, s);
    size_t sb = hg::bf_lowp_scratch_bytes((uint32_t)nq, (uint32_t)b->dim, (uint32_t)c, (uint32_t)b->count);
    b->scratch.reserve(sb);
    hg::LowpReport rep;
    hip_check(hg::launch_bf_lowp(b->data(), precision, dq, (uint32_t)nq, (uint32_t)k, (uint32_t)c, dl, dd,
                                 b->scratch.p, s, stats, variant, (const uint32_t*)b->xstat.p, 1e-6, &rep),
              "reduced-precision brute-force launch");
    b->last = rep;
}

}  // namespace

HG_API int hg_bf_create(int space, size_t dim, size_t max_elements, int device, void** out) {
    return guarded([&] {
        require(out != nullptr, "null output handle");
        require(space == HG_SPACE_L2 || space == HG_SPACE_IP, "Space name must be one of l2, ip, or cosine.");
        require(dim >= 1 && dim <= 65536, "invalid dim");
        require(max_elements < (1ull << 31), "more than 2^31 rows are not supported");
        int ndev = 0;
        hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
        require(device >= 0 && device < ndev, "invalid device");
        DeviceGuard dg(device);
        std::unique_ptr<BfIndex> b(new BfIndex());
        b->device = device;
        b->space = space;
        b->dim = dim;
        b->max_elements = max_elements;
        b->serial = ++g_bf_serial;
        b->vec.reserve(std::max<size_t>(max_elements, 1) * dim * 4);
        b->labels.reserve(std::max<size_t>(max_elements, 1) * 8);
        hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
        *out = b.release();
        return HG_OK;
    });
}

// BruteforceSearch::addPoint (bruteforce.h:64-83) row by row: an existing label overwrites its
// row, a new label takes the next row.  When a new label finds the index full, the rows placed
// before it stay added (as the reference's loop over addPoint leaves them,
// bindings.cpp:811-822) and the call fails with the reference's message.
HG_API int hg_bf_add(void* bv, const float* data, size_t rows, const uint64_t* labels) {
    return guarded([&] {
        require(bv != nullptr, "null index");
        require(rows == 0 || data != nullptr, "null data");
        auto* b = static_cast<BfIndex*>(bv);
        DeviceGuard dg(b->device);
        // rows [0, placed) get a row each; the label map changes only for them
        std::vector<size_t> target(rows);
        std::vector<uint64_t> lab(rows);
        size_t count = b->count, placed = 0;
        bool full = false;
        for (; placed < rows; placed++) {
            const uint64_t l = labels ? labels[placed] : (uint64_t)(b->count + placed);
            auto it = b->label_row.find(l);
            if (it != b->label_row.end()) {
                target[placed] = it->second;
            } else {
                if (count >= b->max_elements) {
                    full = true;
                    break;
                }
                target[placed] = count;
                b->label_row[l] = count;
                count++;
            }
            lab[placed] = l;
        }
        b->host_labels.resize(count);
        for (size_t r2 = 0; r2 < placed; r2++) b->host_labels[target[r2]] = lab[r2];
        b->count = count;
        b->version++;
        b->lp_count = (size_t)-1;  // reduced-precision copy is stale
        // contiguous runs are copied in one transfer
        size_t r = 0;
        while (r < placed) {
            size_t e = r + 1;
            while (e < placed && target[e] == target[e - 1] + 1) e++;
            hip_check(hipMemcpy((char*)b->vec.p + target[r] * b->dim * 4, data + r * b->dim, (e - r) * b->dim * 4,
                                hipMemcpyHostToDevice),
                      "hipMemcpy H2D");
            std::vector<uint64_t> l(lab.begin() + r, lab.begin() + e);
            hip_check(hipMemcpy((char*)b->labels.p + target[r] * 8, l.data(), (e - r) * 8, hipMemcpyHostToDevice),
                      "hipMemcpy H2D");
            r = e;
        }
        if (full) throw Error(HG_ERR_RUNTIME, "The number of elements exceeds the specified limit\n");
        return HG_OK;
    });
}

// Label of every row, in row order.
HG_API int hg_bf_labels(void* bv, uint64_t* labels, size_t capacity) {
    return guarded([&] {
        require(bv != nullptr, "null index");
        auto* b = static_cast<BfIndex*>(bv);
        require(capacity >= b->count && (b->count == 0 || labels), "label buffer too small");
        std::copy(b->host_labels.begin(), b->host_labels.begin() + b->count, labels);
        return HG_OK;
    });
}

HG_API int hg_bf_count(void* bv, size_t* count) {
    return guarded([&] {
        require(bv && count, "null argument");
        *count = static_cast<BfIndex*>(bv)->count;
        return HG_OK;
    });
}

HG_API int hg_bf_free(void* bv) {
    if (!bv) return HG_OK;
    auto* b = static_cast<BfIndex*>(bv);
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(b->device);
    delete b;
    if (prev >= 0) (void)hipSetDevice(prev);
    return HG_OK;
}

HG_API int hg_bf_search(void* bv, const float* queries, size_t nq, size_t k, uint64_t* labels, float* dists) {
    return guarded([&] {
        require(bv && labels && dists, "null argument");
        auto* b = static_cast<BfIndex*>(bv);
        require(k >= 1 && k <= 2048, "k must be in [1, 2048]");
        if (nq == 0) return HG_OK;
        require(queries != nullptr, "null queries");
        DeviceGuard dg(b->device);
        hipStream_t s = b->stream;
        b->q.reserve(nq * b->dim * 4);
        b->ol.reserve(nq * k * 8);
        b->od.reserve(nq * k * 4);
        hip_check(hipMemcpyAsync(b->q.p, queries, nq * b->dim * 4, hipMemcpyHostToDevice, s), "hipMemcpy H2D");
        if (b->count > 0) exact_search(b, (const float*)b->q.p, nq, k, (uint64_t*)b->ol.p, (float*)b->od.p, s);
        hip_check(hipMemcpyAsync(labels, b->ol.p, nq * k * 8, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
        hip_check(hipMemcpyAsync(dists, b->od.p, nq * k * 4, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
        hip_check(hipStreamSynchronize(s), "brute force");
        if (k > b->count) {
            g_err = "Cannot return the results in a contiguous 2D array. There are not enough elements.";
            return HG_ERR_SHORT;
        }
        return HG_OK;
    });
}

HG_API int hg_bf_search_device(void* bv, const float* dq, size_t nq, size_t k, int precision, size_t rerank_c,
                               uint64_t* dl, float* dd, double* stats, void* stream) {
    return guarded([&] {
        require(bv && dq && dl && dd, "null argument");
        auto* b = static_cast<BfIndex*>(bv);
        require(k >= 1 && k <= 2048, "k must be in [1, 2048]");
        require(precision == HG_PREC_FP32 || precision == HG_PREC_BF16 || precision == HG_PREC_FP8,
                "unknown precision");
        require(k <= b->count, "Cannot return the results in a contiguous 2D array. There are not enough elements.");
        if (nq == 0) return HG_OK;
        DeviceGuard dg(b->device);
        hipStream_t s = (hipStream_t)stream;
        if (precision == HG_PREC_FP32) {
            exact_search(b, dq, nq, k, dl, dd, s);
            if (stats) std::fill(stats, stats + 4, 0.0);
            return HG_OK;
        }
        lowp_search(b, dq, nq, k, precision, rerank_c, dl, dd, stats, s, b->tuned[precision]);
        return HG_OK;
    });
}

// BFIndex.knn_query on host buffers at a chosen precision: FP32 is hg_bf_search; BF16 / FP8 run
// the reduced-precision GEMM, re-rank rerank_c candidates in exact FP32 and answer every query
// the error bound cannot certify with the exact path, so the lists equal hg_bf_search's.
HG_API int hg_bf_search_precision(void* bv, const float* queries, size_t nq, size_t k, int precision,
                                  size_t rerank_c, uint64_t* labels, float* dists) {
    if (precision == HG_PREC_FP32) return hg_bf_search(bv, queries, nq, k, labels, dists);
    return guarded([&] {
        require(bv && labels && dists, "null argument");
        auto* b = static_cast<BfIndex*>(bv);
        require(k >= 1 && k <= 2048, "k must be in [1, 2048]");
        require(precision == HG_PREC_BF16 || precision == HG_PREC_FP8, "unknown precision");
        if (nq == 0) return HG_OK;
        require(queries != nullptr, "null queries");
        if (k > b->count) {
            g_err = "Cannot return the results in a contiguous 2D array. There are not enough elements.";
            return HG_ERR_SHORT;
        }
        DeviceGuard dg(b->device);
        hipStream_t s = b->stream;
        b->q.reserve(nq * b->dim * 4);
        b->ol.reserve(nq * k * 8);
        b->od.reserve(nq * k * 4);
        hip_check(hipMemcpyAsync(b->q.p, queries, nq * b->dim * 4, hipMemcpyHostToDevice, s), "hipMemcpy H2D");
        lowp_search(b, (const float*)b->q.p, nq, k, precision, rerank_c, (uint64_t*)b->ol.p, (float*)b->od.p, nullptr,
                    s, b->tuned[precision]);
        hip_check(hipMemcpyAsync(labels, b->ol.p, nq * k * 8, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
        hip_check(hipMemcpyAsync(dists, b->od.p, nq * k * 4, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
        hip_check(hipStreamSynchronize(s), "brute force");
        return HG_OK;
    });
}

HG_API int hg_bf_lowp_report(void* bv, double* out) {
    return guarded([&] {
        require(bv && out, "null argument");
        const hg::LowpReport& r = static_cast<BfIndex*>(bv)->last;
        out[0] = r.exact_overflow;
        out[1] = r.exact_uncertified;
        out[2] = (double)r.near_ties;
        out[3] = (double)r.near_tie_disagree;
        out[4] = (double)r.boundary_near_ties;
        out[5] = (double)r.boundary_disagree;
        out[6] = r.max_bound;
        out[7] = r.certified ? 1.0 : 0.0;
        out[8] = (double)r.reranked_rows;
        return HG_OK;
    });
}

HG_API int hg_bf_fp64_check(void* bv, const float* dq, size_t nq, size_t k, const uint64_t* dl, const float* dd,
                            double rel_tol, double* out, void* stream) {
    return guarded([&] {
        require(bv && dq && dl && dd && out, "null argument");
        auto* b = static_cast<BfIndex*>(bv);
        require(k >= 1 && k <= 2048, "k must be in [1, 2048]");
        if (nq == 0) {
            std::fill(out, out + 4, 0.0);
            return HG_OK;
        }
        DeviceGuard dg(b->device);
        hipStream_t s = (hipStream_t)stream;
        if (b->sorted_version != b->version) {
            std::vector<std::pair<uint64_t, uint32_t>> lr(b->count);
            for (size_t r = 0; r < b->count; r++) lr[r] = std::make_pair(b->host_labels[r], (uint32_t)r);
            std::sort(lr.begin(), lr.end());
            std::vector<uint64_t> sl(b->count);
            std::vector<uint32_t> sr(b->count);
            for (size_t i = 0; i < lr.size(); i++) {
                sl[i] = lr[i].first;
                sr[i] = lr[i].second;
            }
            b->sorted_labels.reserve(std::max<size_t>(sl.size() * 8, 8));
            b->sorted_rows.reserve(std::max<size_t>(sr.size() * 4, 4));
            hip_check(hipMemcpy(b->sorted_labels.p, sl.data(), sl.size() * 8, hipMemcpyHostToDevice), "hipMemcpy H2D");
            hip_check(hipMemcpy(b->sorted_rows.p, sr.data(), sr.size() * 4, hipMemcpyHostToDevice), "hipMemcpy H2D");
            b->sorted_version = b->version;
        }
        b->cand.reserve(64);
        double o[5];
        hip_check(hg::launch_fp64_check(b->data(), dq, (uint32_t)nq, (uint32_t)k, dl, dd, rel_tol,
                                        (const uint64_t*)b->sorted_labels.p, (const uint32_t*)b->sorted_rows.p, o,
                                        b->cand.p, s),
                  "fp64 check launch");
        require(o[4] == 0.0, "a listed label is not in the index");
        std::copy(o, o + 4, out);
        return HG_OK;
    });
}

static_assert(hg::kGemmVariants == HG_GEMM_VARIANTS, "variant count of the C ABI");

// Autotuner of the distance GEMM's tile configuration: runs the reduced-precision search on the
// caller's queries with every variant (each twice, the second timed), keeps the variant with the
// shortest filter GEMM for this index and precision, and reports the times.
HG_API int hg_bf_autotune(void* bv, const float* dq, size_t nq, size_t k, int precision, size_t rerank_c,
                          uint64_t* dl, float* dd, double* times_ms, int* chosen, void* stream) {
    return guarded([&] {
        require(bv && dq && dl && dd, "null argument");
        auto* b = static_cast<BfIndex*>(bv);
        require(precision == HG_PREC_BF16 || precision == HG_PREC_FP8, "autotune needs BF16 or FP8");
        require(k >= 1 && k <= 2048 && k <= b->count && nq > 0, "invalid k or empty batch");
        require(b->dim % 64 == 0 && b->dim <= 8192, "the reduced-precision GEMM needs dim % 64 == 0, dim <= 8192");
        DeviceGuard dg(b->device);
        hipStream_t s = (hipStream_t)stream;
        require(std::max(rerank_c, k) <= 4096, "rerank_c above 4096 is not supported");
        double best = 1e30;
        int best_v = -1;
        for (int v = 0; v < hg::kGemmVariants; v++) {
            double st[4] = {0, 0, 0, 0};
            for (int rep = 0; rep < 2; rep++) lowp_search(b, dq, nq, k, precision, rerank_c, dl, dd, st, s, v);
            if (times_ms) times_ms[v] = st[0];
            if (st[0] < best) {
                best = st[0];
                best_v = v;
            }
        }
        b->tuned[precision] = best_v;
        if (chosen) *chosen = best_v;
        return HG_OK;
    });
}

// BruteforceSearch::removePoint (bruteforce.h:85-104): the last row moves into the removed
// row; an unknown label is ignored, as in the reference.
HG_API int hg_bf_delete(void* bv, uint64_t label) {
    return guarded([&] {
        require(bv != nullptr, "null index");
        auto* b = static_cast<BfIndex*>(bv);
        auto it = b->label_row.find(label);
        if (it == b->label_row.end()) return HG_OK;
        const size_t row = it->second, last = b->count - 1;
        b->label_row.erase(it);
        DeviceGuard dg(b->device);
        if (row != last) {
            const uint64_t moved = b->host_labels[last];
            b->label_row[moved] = row;
            b->host_labels[row] = moved;
            hip_check(hipMemcpy((char*)b->vec.p + row * b->dim * 4, (const char*)b->vec.p + last * b->dim * 4,
                                b->dim * 4, hipMemcpyDeviceToDevice),
                      "hipMemcpy D2D");
            hip_check(hipMemcpy((char*)b->labels.p + row * 8, &moved, 8, hipMemcpyHostToDevice), "hipMemcpy H2D");
        }
        b->host_labels.pop_back();
        b->count--;
        b->version++;
        b->lp_count = (size_t)-1;
        return HG_OK;
    });
}

// BruteforceSearch::saveIndex (bruteforce.h:128-141): max_elements, bytes per element and the
// element count (size_t each), then max_elements records of [dim f32 | u64 label]; records past
// the count are zero.
HG_API int hg_bf_save(void* bv, const char* path) {
    return guarded([&] {
        require(bv && path, "null argument");
        auto* b = static_cast<BfIndex*>(bv);
        DeviceGuard dg(b->device);
        std::ofstream out(path, std::ios::binary);
        require((bool)out, std::string("Cannot open file ") + path);
        const size_t per = b->dim * 4 + 8;
        const size_t hdr[3] = {b->max_elements, per, b->count};
        out.write((const char*)hdr, sizeof(hdr));
        std::vector<float> vec(b->count * b->dim);
        if (b->count)
            hip_check(hipMemcpy(vec.data(), b->vec.p, vec.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy D2H");
        std::vector<char> rec(per);
        for (size_t r = 0; r < b->count; r++) {
            memcpy(rec.data(), &vec[r * b->dim], b->dim * 4);
            memcpy(rec.data() + b->dim * 4, &b->host_labels[r], 8);
            out.write(rec.data(), per);
        }
        std::fill(rec.begin(), rec.end(), 0);
        for (size_t r = b->count; r < b->max_elements; r++) out.write(rec.data(), per);
        require((bool)out, "write failed");
        return HG_OK;
    });
}

// BruteforceSearch::loadIndex (bruteforce.h:144-161) into a new GPU index; the label -> row
// map is rebuilt from the records (the reference leaves it empty after a load).
HG_API int hg_bf_load(int space, size_t dim, const char* path, int device, void** out) {
    return guarded([&] {
        require(path && out, "null argument");
        std::ifstream in(path, std::ios::binary);
        require((bool)in, std::string("Cannot open file ") + path);
        in.seekg(0, in.end);
        const size_t file_bytes = (size_t)in.tellg();
        in.seekg(0, in.beg);
        size_t hdr[3] = {0, 0, 0};
        in.read((char*)hdr, sizeof(hdr));
        const size_t per = dim * 4 + 8;
        // the stored count must fit the file (no allocation sized by a damaged header)
        require((bool)in && hdr[2] <= hdr[0] && hdr[2] <= (file_bytes - sizeof(hdr)) / per,
                "Index seems to be corrupted or unsupported");
        require(hdr[1] == per, "Index seems to be corrupted or unsupported");
        void* h = nullptr;
        int rc = hg_bf_create(space, dim, hdr[0], device, &h);
        if (rc != HG_OK) return rc;
        std::unique_ptr<BfIndex> b(static_cast<BfIndex*>(h));
        std::vector<float> vec(hdr[2] * dim);
        std::vector<uint64_t> lab(hdr[2]);
        std::vector<char> rec(per);
        for (size_t r = 0; r < hdr[2]; r++) {
            in.read(rec.data(), per);
            require((bool)in, "Index seems to be corrupted or unsupported");
            memcpy(&vec[r * dim], rec.data(), dim * 4);
            memcpy(&lab[r], rec.data() + dim * 4, 8);
        }
        rc = hg_bf_add(b.get(), vec.data(), hdr[2], lab.data());
        if (rc != HG_OK) return rc;
        *out = b.release();
        return HG_OK;
    });
}

HG_API int hg_bf_filter_create(void* bv, const uint64_t* allowed_labels, size_t n_allowed, void** out) {
    return guarded([&] {
        require(bv && out && (n_allowed == 0 || allowed_labels), "null argument");
        auto* b = static_cast<BfIndex*>(bv);
        std::vector<uint32_t> bits((b->count + 31) / 32 + 1, 0u);
        std::unique_ptr<BfFilter> f(new BfFilter());
        for (size_t i = 0; i < n_allowed; i++) {
            auto it = b->label_row.find(allowed_labels[i]);
            if (it == b->label_row.end()) continue;
            const size_t r = it->second;
            if (!((bits[r >> 5] >> (r & 31)) & 1u)) f->allowed++;
            bits[r >> 5] |= 1u << (r & 31);
        }
        DeviceGuard dg(b->device);
        f->serial = b->serial;
        f->version = b->version;
        f->device = b->device;
        f->bits.reserve(bits.size() * 4);
        hip_check(hipMemcpy(f->bits.p, bits.data(), bits.size() * 4, hipMemcpyHostToDevice), "hipMemcpy H2D");
        *out = f.release();
        return HG_OK;
    });
}

HG_API int hg_bf_filter_free(void* f) {
    if (!f) return HG_OK;
    auto* bf = static_cast<BfFilter*>(f);
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(bf->device);
    delete bf;
    if (prev >= 0) (void)hipSetDevice(prev);
    return HG_OK;
}

// BFIndex.knn_query(..., filter=) on host buffers (exact FP32): the k smallest (distance, label)
// pairs among the allowed rows; HG_ERR_SHORT when fewer than k rows are allowed.
HG_API int hg_bf_search_filtered(void* bv, const float* queries, size_t nq, size_t k, void* filter,
                                 uint64_t* labels, float* dists) {
    return guarded([&] {
        require(bv && labels && dists, "null argument");
        auto* b = static_cast<BfIndex*>(bv);
        require(k >= 1 && k <= 2048, "k must be in [1, 2048]");
        BfFilter* f = as_bf_filter(b, filter);
        if (nq == 0) return HG_OK;
        require(queries != nullptr, "null queries");
        DeviceGuard dg(b->device);
        hipStream_t s = b->stream;
        b->q.reserve(nq * b->dim * 4);
        b->ol.reserve(nq * k * 8);
        b->od.reserve(nq * k * 4);
        hip_check(hipMemcpyAsync(b->q.p, queries, nq * b->dim * 4, hipMemcpyHostToDevice, s), "hipMemcpy H2D");
        if (b->count > 0)
            exact_search(b, (const float*)b->q.p, nq, k, (uint64_t*)b->ol.p, (float*)b->od.p, s,
                         f ? (const uint32_t*)f->bits.p : nullptr);
        hip_check(hipMemcpyAsync(labels, b->ol.p, nq * k * 8, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
        hip_check(hipMemcpyAsync(dists, b->od.p, nq * k * 4, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
        hip_check(hipStreamSynchronize(s), "brute force");
        if (k > (f ? f->allowed : b->count)) {
            g_err = "Cannot return the results in a contiguous 2D array. There are not enough elements.";
            return HG_ERR_SHORT;
        }
        return HG_OK;
    });
}

// Device-resident filtered exact search (FP32), asynchronous on stream.
HG_API int hg_bf_search_device_filtered(void* bv, const float* dq, size_t nq, size_t k, void* filter, uint64_t* dl,
                                        float* dd, void* stream) {
    return guarded([&] {
        require(bv && dq && dl && dd, "null argument");
        auto* b = static_cast<BfIndex*>(bv);
        require(k >= 1 && k <= 2048, "k must be in [1, 2048]");
        BfFilter* f = as_bf_filter(b, filter);
        require(k <= (f ? f->allowed : b->count),
                "Cannot return the results in a contiguous 2D array. There are not enough elements.");
        if (nq == 0) return HG_OK;
        DeviceGuard dg(b->device);
        exact_search(b, dq, nq, k, dl, dd, (hipStream_t)stream, f ? (const uint32_t*)f->bits.p : nullptr);
        return HG_OK;
    });
}

HG_API int hg_bf_gemm_variant(int variant) {
    return guarded([&] {
        require(variant >= 0 && variant < hg::kGemmVariants, "GEMM variant must be in [0, 4]");
        hg::set_gemm_variant(variant);
        return HG_OK;
    });
}
