// Drives most of the C ABI through the GPU paths: the graph search (host and device buffers,
// per-query ef and k, both row formats, hop caps, filters, the allocation-free entry point, range
// and multi-vector searches, hipGraph capture and replay, delete marks, the stop model, the
// layout cache), the GPU construction and the brute force at every precision, including the
// error paths (beam beyond the LDS, ef beyond the limit, a stale graph, a filter of another
// index, a full index).  Built with the library's host code under AddressSanitizer by
// tests/cpp/host_asan_gpu.sh; any sanitizer report or failed call exits non-zero.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>
#include <unistd.h>

#include "hnswgpu.h"

static int fails = 0;
#define OK(call)                                                                              \
    do {                                                                                      \
        int rc_ = (call);                                                                     \
        if (rc_ != HG_OK) {                                                                   \
            std::printf("FAIL %s -> %d: %s (line %d)\n", #call, rc_, hg_last_error(), __LINE__); \
            fails++;                                                                          \
        }                                                                                     \
    } while (0)
#define ERR(call)                                                                              \
    do {                                                                                       \
        int rc_ = (call);                                                                      \
        if (rc_ == HG_OK) {                                                                    \
            std::printf("FAIL expected an error from %s (line %d)\n", #call, __LINE__);        \
            fails++;                                                                           \
        } else                                                                                 \
            std::printf("  expected error: %s\n", hg_last_error());                            \
    } while (0)

template <typename T>
struct Dev {
    T* p = nullptr;
    size_t n = 0;
    explicit Dev(size_t n_) : n(n_) { OK(hg_malloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T))); }
    Dev(const std::vector<T>& v) : Dev(v.size()) { OK(hg_memcpy_h2d(p, v.data(), v.size() * sizeof(T), nullptr)); }
    std::vector<T> get() const {
        std::vector<T> v(n);
        OK(hg_memcpy_d2h(v.data(), p, n * sizeof(T), nullptr));
        return v;
    }
    ~Dev() { hg_free(p); }
};

static std::vector<float> rows(size_t n, size_t dim, unsigned seed, bool integers) {
    std::mt19937 g(seed);
    std::gamma_distribution<float> ga(0.5f, 40.f);
    std::normal_distribution<float> nd(0.f, 1.f);
    std::vector<float> v(n * dim);
    for (auto& x : v) x = integers ? std::min(255.f, std::floor(ga(g))) : nd(g);
    return v;
}

int main() {
    setvbuf(stdout, nullptr, _IONBF, 0);
    const std::string tmp = std::getenv("TMPDIR") ? std::getenv("TMPDIR") : "/tmp";
    void* stream = nullptr;
    OK(hg_stream_create(&stream));
    // ---------------- HNSW, L2, dim 128, integer rows (8-bit copy) ----------------
    const size_t n = 20000, dim = 128, nq = 300, k = 10;
    auto data = rows(n, dim, 1, true), q = rows(nq, dim, 2, true);
    std::vector<uint64_t> labels(n);
    for (size_t i = 0; i < n; i++) labels[i] = 1000 + i;
    void* h = nullptr;
    OK(hg_hnsw_create(0, dim, n, 16, 100, 100, &h));
    OK(hg_hnsw_add_items(h, data.data(), n, labels.data(), 4));
    std::printf("built the index\n");
    void* g = nullptr;
    OK(hg_gpu_index_create(h, 0, &g));
    int code = -1;
    OK(hg_gpu_vector_code(g, &code));
    std::printf("vector code %d\n", code);
    std::vector<uint64_t> L(nq * k);
    std::vector<float> D(nq * k);
    std::vector<int32_t> C(nq * 4);
    OK(hg_gpu_search(g, q.data(), nq, k, 64, L.data(), D.data(), C.data()));
    Dev<float> dq(q);
    Dev<uint64_t> dl(nq * 16);
    Dev<float> dd(nq * 16);
    Dev<int32_t> dc(nq * 4);
    std::vector<int32_t> efq(nq), kq(nq), caps(nq);
    for (size_t i = 0; i < nq; i++) {
        efq[i] = 8 + (int32_t)(i * 37 % 292);
        kq[i] = 1 + (int32_t)(i % 16);
        caps[i] = (int32_t)(i % 60);
    }
    Dev<int32_t> def(efq), dk(kq), dcap(caps);
    OK(hg_gpu_search_device(g, dq.p, nq, def.p, dk.p, 300, 16, 16, dl.p, dd.p, dc.p, nullptr, stream));
    OK(hg_gpu_set_vector_code(g, HG_VEC_F32));
    OK(hg_gpu_search_device(g, dq.p, nq, def.p, dk.p, 300, 16, 16, dl.p, dd.p, dc.p, nullptr, stream));
    OK(hg_gpu_set_vector_code(g, HG_VEC_AUTO));
    OK(hg_gpu_search_device_hopcap(g, dq.p, nq, dcap.p, 100, k, dl.p, dd.p, dc.p, stream));
    OK(hg_stream_synchronize(stream));
    // plans
    int32_t plan[5];
    OK(hg_gpu_search_plan(g, 200, 10, nq, plan));
    OK(hg_gpu_autotune(g, dq.p, nq, 200, 10, stream));
    OK(hg_gpu_set_plan(g, 200, 10, 7, 12));
    ERR(hg_gpu_set_plan(g, 4000, 10, 15, 12));  // 128 KiB of hash + 32 KiB of beam
    ERR(hg_gpu_search(g, q.data(), nq, k, 5000, L.data(), D.data(), nullptr));
    // filters, status, allocation-free
    std::vector<uint64_t> allowed;
    for (size_t i = 0; i < n; i += 3) allowed.push_back(labels[i]);
    void* f = nullptr;
    OK(hg_gpu_filter_create(g, allowed.data(), allowed.size(), &f));
    size_t na = 0;
    OK(hg_gpu_filter_count(f, &na));
    Dev<uint32_t> ds(1);
    OK(hg_gpu_search_device_filtered(g, dq.p, nq, def.p, dk.p, 300, 16, 16, dl.p, dd.p, dc.p, f, ds.p, stream));
    OK(hg_gpu_search_device_noalloc(g, dq.p, nq, nullptr, nullptr, 64, k, k, dl.p, dd.p, dc.p, nullptr, ds.p, stream));
    OK(hg_gpu_search_device_noalloc(g, dq.p, nq / 2, nullptr, nullptr, 64, k, k, dl.p, dd.p, dc.p, f, ds.p, stream));
    OK(hg_stream_synchronize(stream));
    uint32_t st = 0, sp = 0;
    OK(hg_gpu_search_status(g, &st, &sp));
    // range and multi-vector
    std::vector<uint32_t> cnt(nq);
    std::vector<uint64_t> RL(nq * 400);
    std::vector<float> RD(nq * 400);
    OK(hg_gpu_search_epsilon(g, q.data(), nq, 90000.f, 50, 400, f, 400, RL.data(), RD.data(), cnt.data()));
    Dev<uint64_t> drl(nq * 400);
    Dev<float> drd(nq * 400);
    Dev<uint32_t> dcnt(nq);
    OK(hg_gpu_search_epsilon_device(g, dq.p, nq, 90000.f, 20, 200, nullptr, 400, drl.p, drd.p, dcnt.p, dc.p, ds.p,
                                    stream));
    std::vector<uint32_t> docs(n);
    for (size_t i = 0; i < n; i++) docs[i] = (uint32_t)(i % 997);
    void* dh = nullptr;
    OK(hg_gpu_docs_create(g, labels.data(), docs.data(), n, &dh));
    OK(hg_gpu_search_multivector(g, q.data(), nq, dh, 10, 40, nullptr, 400, RL.data(), RD.data(), cnt.data()));
    OK(hg_gpu_search_multivector_device(g, dq.p, nq, dh, 10, 40, f, 400, drl.p, drd.p, dcnt.p, dc.p, ds.p, 0, stream));
    OK(hg_stream_synchronize(stream));
    // graph capture and replays
    void* graph = nullptr;
    OK(hg_graph_capture_search(g, dq.p, nq, def.p, dk.p, 300, 16, 16, dl.p, dd.p, nullptr, nullptr, stream, &graph));
    for (int r = 0; r < 3; r++) OK(hg_graph_launch(graph, stream));
    OK(hg_stream_synchronize(stream));
    // stop model
    void* sm = nullptr;
    OK(hg_stop_model_create(12, 64, 0, &sm));
    size_t np = 0;
    OK(hg_stop_model_param_count(sm, &np));
    std::vector<float> params(np, 0.01f);
    params[np - 1] = -30.f;
    OK(hg_stop_model_set_params(sm, params.data(), np));
    std::vector<float> mean(12, 0.f), inv(12, 1.f);
    OK(hg_stop_model_set_norm(sm, mean.data(), inv.data(), 12));
    OK(hg_stop_model_config(sm, 16, 0.9f, 16));
    OK(hg_gpu_search_device(g, dq.p, nq, nullptr, nullptr, 200, k, k, dl.p, dd.p, dc.p, sm, stream));
    OK(hg_stop_model_config_extend(sm, 400, 0.3f));
    OK(hg_gpu_search_device(g, dq.p, nq, nullptr, nullptr, 200, k, k, dl.p, dd.p, dc.p, sm, stream));
    OK(hg_stream_synchronize(stream));
    // delete marks: the captured graph must refuse to replay after a refresh
    OK(hg_hnsw_mark_deleted(h, labels[5]));
    OK(hg_hnsw_mark_deleted(h, labels[77]));
    OK(hg_gpu_refresh_deleted(g, h));
    ERR(hg_graph_launch(graph, stream));
    OK(hg_graph_free(graph));
    OK(hg_gpu_search(g, q.data(), nq, k, 64, L.data(), D.data(), nullptr));
    for (auto l : L)
        if (l == labels[5] || l == labels[77]) {
            std::printf("FAIL deleted label returned\n");
            fails++;
        }
    // layout cache
    const std::string lp = tmp + "/abi_layout.bin";
    OK(hg_gpu_index_save(g, lp.c_str()));
    void* g2 = nullptr;
    OK(hg_gpu_index_load(lp.c_str(), 0, &g2));
    OK(hg_gpu_search(g2, q.data(), nq, k, 64, L.data(), D.data(), nullptr));
    ERR(hg_gpu_search_device_filtered(g2, dq.p, nq, nullptr, nullptr, 64, k, k, dl.p, dd.p, nullptr, f, ds.p, stream));
    OK(hg_gpu_index_free(g2));
    std::remove(lp.c_str());
    OK(hg_gpu_docs_free(dh));
    OK(hg_gpu_filter_free(f));
    OK(hg_stop_model_free(sm));
    OK(hg_gpu_index_free(g));
    OK(hg_hnsw_free(h));
    // ---------------- GPU construction, IP, dim 96, float rows ----------------
    {
        const size_t n2 = 12000, d2 = 96;
        auto x = rows(n2, d2, 3, false), q2 = rows(100, d2, 4, false);
        void* h2 = nullptr;
        OK(hg_hnsw_create(1, d2, n2, 16, 100, 100, &h2));
        OK(hg_hnsw_build_gpu(h2, x.data(), n2, nullptr, 4, 0));
        void* g3 = nullptr;
        OK(hg_gpu_index_create(h2, 0, &g3));
        std::vector<uint64_t> L2(100 * 10);
        std::vector<float> D2(100 * 10);
        OK(hg_gpu_search(g3, q2.data(), 100, 10, 50, L2.data(), D2.data(), nullptr));
        OK(hg_gpu_index_free(g3));
        OK(hg_hnsw_free(h2));
    }
    // ---------------- brute force, IP, dim 128 ----------------
    {
        const size_t n3 = 30000, d3 = 128, q3n = 200, k3 = 20;
        auto x = rows(n3, d3, 5, false), q3 = rows(q3n, d3, 6, false);
        std::vector<uint64_t> lab(n3);
        for (size_t i = 0; i < n3; i++) lab[i] = 7 * i + 3;
        void* b = nullptr;
        OK(hg_bf_create(1, d3, n3, 0, &b));
        OK(hg_bf_add(b, x.data(), n3, lab.data()));
        ERR(hg_bf_add(b, x.data(), 1, nullptr));  // past the capacity
        std::vector<uint64_t> BL(q3n * k3);
        std::vector<float> BD(q3n * k3);
        OK(hg_bf_search(b, q3.data(), q3n, k3, BL.data(), BD.data()));
        Dev<float> dq3(q3);
        Dev<uint64_t> dbl(q3n * k3);
        Dev<float> dbd(q3n * k3);
        double stats[4];
        for (int prec : {HG_PREC_FP32, HG_PREC_BF16, HG_PREC_FP8})
            OK(hg_bf_search_device(b, dq3.p, q3n, k3, prec, 512, dbl.p, dbd.p, stats, stream));
        double times[HG_GEMM_VARIANTS];
        int chosen = -1;
        OK(hg_bf_autotune(b, dq3.p, q3n, k3, HG_PREC_BF16, 512, dbl.p, dbd.p, times, &chosen, stream));
        double rep[9];
        OK(hg_bf_lowp_report(b, rep));
        double fp64[5];
        OK(hg_bf_fp64_check(b, dq3.p, q3n, k3, dbl.p, dbd.p, 1e-6, fp64, stream));
        OK(hg_bf_search_precision(b, q3.data(), q3n, k3, HG_PREC_FP8, 256, BL.data(), BD.data()));
        std::vector<uint64_t> al;
        for (size_t i = 0; i < n3; i += 5) al.push_back(lab[i]);
        void* bfil = nullptr;
        OK(hg_bf_filter_create(b, al.data(), al.size(), &bfil));
        OK(hg_bf_search_filtered(b, q3.data(), q3n, k3, bfil, BL.data(), BD.data()));
        OK(hg_bf_search_device_filtered(b, dq3.p, q3n, k3, bfil, dbl.p, dbd.p, stream));
        OK(hg_stream_synchronize(stream));
        OK(hg_bf_filter_free(bfil));
        OK(hg_bf_delete(b, lab[10]));
        OK(hg_bf_delete(b, 999999999ull));  // unknown label: ignored
        const std::string bp = tmp + "/abi_bf.bin";
        OK(hg_bf_save(b, bp.c_str()));
        void* b2 = nullptr;
        OK(hg_bf_load(1, d3, bp.c_str(), 0, &b2));
        size_t c2 = 0;
        OK(hg_bf_count(b2, &c2));
        if (c2 != n3 - 1) {
            std::printf("FAIL loaded count %zu\n", c2);
            fails++;
        }
        OK(hg_bf_search(b2, q3.data(), q3n, k3, BL.data(), BD.data()));
        OK(hg_bf_free(b2));
        std::remove(bp.c_str());
        OK(hg_bf_free(b));
    }
    OK(hg_stream_destroy(stream));
    std::printf("abi drive: %s\n", fails ? "FAILED" : "ok");
    std::fflush(stdout);
    _exit(fails ? 1 : 0);  // skip the runtime's finalizers (the sanitizer's allocator checks trip in them)
}
