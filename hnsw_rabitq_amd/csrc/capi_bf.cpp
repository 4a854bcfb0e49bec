// C ABI: brute-force index (BruteforceSearch / BFIndex on the GPU).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
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
    int lp_kind = 0;  // precision of vec_lp (0 = not built)
    int tuned[3] = {-1, -1, -1};  // GEMM variant chosen by the autotuner, per precision
    size_t lp_count = 0;
    hipStream_t stream = nullptr;
    std::unordered_map<uint64_t, size_t> label_row;
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
        d.n = (uint32_t)count;
        d.dim = (uint32_t)dim;
        d.space = space;
        return d;
    }
};

// Exact FP32 search with device buffers on stream s.
void exact_search(BfIndex* b, const float* dq, size_t nq, size_t k, uint64_t* dl, float* dd, hipStream_t s) {
    hg::BfData d = b->data();
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
        b->vec.reserve(std::max<size_t>(max_elements, 1) * dim * 4);
        b->labels.reserve(std::max<size_t>(max_elements, 1) * 8);
        hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
        *out = b.release();
        return HG_OK;
    });
}

// BruteforceSearch::addPoint (bruteforce.h:64-83): an existing label overwrites its row,
// a new label takes the next row; capacity errors use the reference's message.
HG_API int hg_bf_add(void* bv, const float* data, size_t rows, const uint64_t* labels) {
    return guarded([&] {
        require(bv != nullptr, "null index");
        require(rows == 0 || data != nullptr, "null data");
        auto* b = static_cast<BfIndex*>(bv);
        DeviceGuard dg(b->device);
        std::vector<size_t> target(rows);
        std::vector<uint64_t> lab(rows);
        size_t count = b->count;
        for (size_t r = 0; r < rows; r++) {
            uint64_t l = labels ? labels[r] : (uint64_t)(b->count + r);
            auto it = b->label_row.find(l);
            if (it != b->label_row.end()) {
                target[r] = it->second;
            } else {
                if (count >= b->max_elements)
                    throw Error(HG_ERR_RUNTIME, "The number of elements exceeds the specified limit\n");
                target[r] = count;
                b->label_row[l] = count;
                count++;
            }
            lab[r] = l;
        }
        // contiguous runs are copied in one transfer
        size_t r = 0;
        while (r < rows) {
            size_t e = r + 1;
            while (e < rows && target[e] == target[e - 1] + 1) e++;
            hip_check(hipMemcpy((char*)b->vec.p + target[r] * b->dim * 4, data + r * b->dim, (e - r) * b->dim * 4,
                                hipMemcpyHostToDevice),
                      "hipMemcpy H2D");
            std::vector<uint64_t> l(lab.begin() + r, lab.begin() + e);
            hip_check(hipMemcpy((char*)b->labels.p + target[r] * 8, l.data(), (e - r) * 8, hipMemcpyHostToDevice),
                      "hipMemcpy H2D");
            r = e;
        }
        b->count = count;
        b->lp_count = (size_t)-1;  // reduced-precision copy is stale
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
        require(b->dim % 64 == 0, "the reduced-precision GEMM needs dim to be a multiple of 64");
        require(b->dim <= 8192, "the reduced-precision path supports dim up to 8192");
        size_t c = std::max(rerank_c, k);
        require(c <= 4096, "rerank_c above 4096 is not supported");
        c = std::min(c, b->count);
        ensure_lowp(b, precision, s);
        size_t sb = hg::bf_lowp_scratch_bytes((uint32_t)nq, (uint32_t)b->dim, (uint32_t)c, (uint32_t)b->count);
        b->scratch.reserve(sb);
        hip_check(hg::launch_bf_lowp(b->data(), precision, dq, (uint32_t)nq, (uint32_t)k, (uint32_t)c, dl, dd,
                                     b->scratch.p, s, stats, b->tuned[precision]),
                  "reduced-precision brute-force launch");
        return HG_OK;
    });
}

HG_API int hg_bf_fp64_check(void* bv, const float* dq, size_t nq, size_t k, const uint64_t* dl, const float* dd,
                            double rel_tol, double* out, void* stream) {
    return guarded([&] {
        require(bv && dq && dl && dd && out, "null argument");
        auto* b = static_cast<BfIndex*>(bv);
        DeviceGuard dg(b->device);
        b->cand.reserve(std::max<size_t>(nq * k * 8, 64));
        hip_check(hg::launch_fp64_check(b->data(), dq, (uint32_t)nq, (uint32_t)k, dl, dd, rel_tol, out, b->cand.p,
                                        (hipStream_t)stream),
                  "fp64 check launch");
        return HG_OK;
    });
}

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
        size_t c = std::min(std::max(rerank_c, k), b->count);
        require(c <= 4096, "rerank_c above 4096 is not supported");
        ensure_lowp(b, precision, s);
        b->scratch.reserve(hg::bf_lowp_scratch_bytes((uint32_t)nq, (uint32_t)b->dim, (uint32_t)c, (uint32_t)b->count));
        double best = 1e30;
        int best_v = -1;
        for (int v = 0; v < hg::kGemmVariants; v++) {
            double st[4] = {0, 0, 0, 0};
            for (int rep = 0; rep < 2; rep++)
                hip_check(hg::launch_bf_lowp(b->data(), precision, dq, (uint32_t)nq, (uint32_t)k, (uint32_t)c, dl, dd,
                                             b->scratch.p, s, st, v),
                          "reduced-precision brute-force launch");
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

HG_API int hg_bf_gemm_variant(int variant) {
    return guarded([&] {
        require(variant >= 0 && variant < hg::kGemmVariants, "GEMM variant must be in [0, 4]");
        hg::set_gemm_variant(variant);
        return HG_OK;
    });
}
