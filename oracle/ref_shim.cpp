// Driver around the reference hnswlib headers, compiled from the sources where they lie
// (reference: hnswlib/hnswlib.h, hnswalg.h, bruteforce.h, space_l2.h, space_ip.h).
// Test infrastructure only: it produces golden vectors and cross-checks the CPU restatement
// in oracle.cpp.  Nothing here is shipped or called by the product library.
//
// The flags mirror the Python extension build of the reference (setup.py:76-103:
// -O3 -march=native -fopenmp -std=c++14), so the distance bits are those the reference's
// knn_query produces on an AVX-512 host.
#include "hnswlib/hnswlib.h"

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {
struct RefSpace {
    hnswlib::SpaceInterface<float>* sp = nullptr;
    RefSpace(int space, size_t dim) {
        if (space == 0) sp = new hnswlib::L2Space(dim);
        else sp = new hnswlib::InnerProductSpace(dim);
    }
    ~RefSpace() { delete sp; }
};

struct RefHnsw {
    RefSpace space;
    hnswlib::HierarchicalNSW<float>* alg = nullptr;
    RefHnsw(int s, size_t dim) : space(s, dim) {}
    ~RefHnsw() { delete alg; }
};

std::string g_err;
}  // namespace

REF_API const char* ref_last_error() { return g_err.c_str(); }

// Distance exactly as the reference space computes it (space 0 = l2, 1 = ip).
REF_API float ref_distance(int space, size_t dim, const float* a, const float* b) {
    RefSpace s(space, dim);
    return s.sp->get_dist_func()(a, b, s.sp->get_dist_func_param());
}

REF_API void ref_distance_batch(int space, size_t dim, const float* a, const float* b, size_t n, float* out) {
    RefSpace s(space, dim);
    auto f = s.sp->get_dist_func();
    void* p = s.sp->get_dist_func_param();
    for (size_t i = 0; i < n; i++) out[i] = f(a + i * dim, b + i * dim, p);
}

// Build single-threaded in insertion order (labels = 0..n-1), as Index.add_items does for the
// first point and then ParallelFor with one thread (bindings.cpp:251-304).
REF_API void* ref_hnsw_build(int space, size_t dim, const float* data, size_t n, size_t M, size_t efc,
                             size_t seed) {
    try {
        RefHnsw* h = new RefHnsw(space, dim);
        h->alg = new hnswlib::HierarchicalNSW<float>(h->space.sp, n, M, efc, seed);
        for (size_t i = 0; i < n; i++) h->alg->addPoint(data + i * dim, (hnswlib::labeltype)i);
        return h;
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

REF_API void* ref_hnsw_load(int space, size_t dim, const char* path) {
    try {
        RefHnsw* h = new RefHnsw(space, dim);
        h->alg = new hnswlib::HierarchicalNSW<float>(h->space.sp, std::string(path));
        return h;
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

REF_API int ref_hnsw_save(void* hv, const char* path) {
    try {
        static_cast<RefHnsw*>(hv)->alg->saveIndex(std::string(path));
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

REF_API void ref_hnsw_free(void* hv) { delete static_cast<RefHnsw*>(hv); }

REF_API int ref_hnsw_maxlevel(void* hv) { return static_cast<RefHnsw*>(hv)->alg->maxlevel_; }
REF_API unsigned ref_hnsw_entry(void* hv) { return static_cast<RefHnsw*>(hv)->alg->enterpoint_node_; }

// knn_query semantics (bindings.cpp:612-697): ef_ set globally, searchKnn per row, rows written
// closest-first.  Returns number of rows that produced fewer than k results (those rows are
// left with label = UINT64_MAX, dist = +inf).
REF_API int ref_hnsw_search(void* hv, const float* q, size_t nq, size_t k, size_t ef, uint64_t* labels,
                            float* dists) {
    RefHnsw* h = static_cast<RefHnsw*>(hv);
    h->alg->setEf(ef);
    size_t dim = *(size_t*)h->space.sp->get_dist_func_param();
    int short_rows = 0;
    for (size_t r = 0; r < nq; r++) {
        auto res = h->alg->searchKnn(q + r * dim, k);
        for (size_t i = 0; i < k; i++) {
            labels[r * k + i] = UINT64_MAX;
            dists[r * k + i] = __builtin_inff();
        }
        if (res.size() != k) short_rows++;
        for (int i = (int)res.size() - 1; i >= 0; i--) {
            labels[r * k + i] = res.top().second;
            dists[r * k + i] = res.top().first;
            res.pop();
        }
    }
    return short_rows;
}

// markDelete (hnswalg.h:852-868); returns -1 with the reference's message on error.
REF_API int ref_hnsw_mark_deleted(void* hv, uint64_t label) {
    try {
        static_cast<RefHnsw*>(hv)->alg->markDelete((hnswlib::labeltype)label);
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

namespace {
// A label filter as a caller of searchKnn would pass it (hnswlib.h:128-132).
struct LabelMask : public hnswlib::BaseFilterFunctor {
    const uint8_t* mask;
    size_t n;
    LabelMask(const uint8_t* m, size_t nn) : mask(m), n(nn) {}
    bool operator()(hnswlib::labeltype l) override { return l < n && mask[l] != 0; }
};
}  // namespace

// ref_hnsw_search with a filter over labels (allowed[label] != 0; null = no filter): searchKnn
// takes the non-bare-bone branch whenever a filter is given or elements are deleted.
REF_API int ref_hnsw_search_filtered(void* hv, const float* q, size_t nq, size_t k, size_t ef,
                                     const uint8_t* allowed, size_t n_labels, uint64_t* labels, float* dists) {
    RefHnsw* h = static_cast<RefHnsw*>(hv);
    h->alg->setEf(ef);
    size_t dim = *(size_t*)h->space.sp->get_dist_func_param();
    LabelMask f(allowed, n_labels);
    int short_rows = 0;
    for (size_t r = 0; r < nq; r++) {
        auto res = h->alg->searchKnn(q + r * dim, k, allowed ? &f : nullptr);
        for (size_t i = 0; i < k; i++) {
            labels[r * k + i] = UINT64_MAX;
            dists[r * k + i] = __builtin_inff();
        }
        if (res.size() != k) short_rows++;
        for (int i = (int)res.size() - 1; i >= 0; i--) {
            labels[r * k + i] = res.top().second;
            dists[r * k + i] = res.top().first;
            res.pop();
        }
    }
    return short_rows;
}

// searchStopConditionClosest with EpsilonSearchStopCondition (hnswalg.h:1326-1377,
// stop_condition.h:207-275), optionally filtered.  Row r holds min(counts[r], out_stride) of its
// results in the reference's order; the second member of each pair is what the reference puts
// there (the internal id of the result).
REF_API int ref_hnsw_search_epsilon(void* hv, const float* q, size_t nq, float eps, size_t min_c, size_t max_c,
                                    const uint8_t* allowed, size_t n_labels, size_t out_stride, uint64_t* labels,
                                    float* dists, uint32_t* counts) {
    try {
        RefHnsw* h = static_cast<RefHnsw*>(hv);
        size_t dim = *(size_t*)h->space.sp->get_dist_func_param();
        LabelMask f(allowed, n_labels);
        for (size_t r = 0; r < nq; r++) {
            hnswlib::EpsilonSearchStopCondition<float> sc(eps, min_c, max_c);
            auto res = h->alg->searchStopConditionClosest(q + r * dim, sc, allowed ? &f : nullptr);
            counts[r] = (uint32_t)res.size();
            for (size_t i = 0; i < out_stride; i++) {
                labels[r * out_stride + i] = i < res.size() ? (uint64_t)res[i].second : UINT64_MAX;
                dists[r * out_stride + i] = i < res.size() ? res[i].first : __builtin_inff();
            }
        }
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// A multi-vector index (MultiVectorL2Space<unsigned>, stop_condition.h:19-76): each element is
// its vector followed by its document id; built single-threaded in insertion order (labels =
// 0..n-1).  The graph equals the plain L2 build's (the distance reads the vector only).
namespace {
struct RefMv {
    hnswlib::MultiVectorL2Space<unsigned> space;
    hnswlib::HierarchicalNSW<float>* alg = nullptr;
    size_t dim;
    explicit RefMv(size_t d) : space(d), dim(d) {}
    ~RefMv() { delete alg; }
};
}  // namespace

REF_API void* ref_mv_build(size_t dim, const float* data, const uint32_t* docs, size_t n, size_t M, size_t efc,
                           size_t seed) {
    try {
        RefMv* h = new RefMv(dim);
        h->alg = new hnswlib::HierarchicalNSW<float>(&h->space, n, M, efc, seed);
        std::vector<char> point(h->space.get_data_size());
        for (size_t i = 0; i < n; i++) {
            memcpy(point.data(), data + i * dim, dim * 4);
            h->space.set_doc_id(point.data(), docs[i]);
            h->alg->addPoint(point.data(), (hnswlib::labeltype)i);
        }
        return h;
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

REF_API void ref_mv_free(void* hv) { delete static_cast<RefMv*>(hv); }

// searchStopConditionClosest with MultiVectorSearchStopCondition(space, num_docs, ef_collection)
// (hnswalg.h:1326-1377, stop_condition.h:82-200); rows as ref_hnsw_search_epsilon.
REF_API int ref_mv_search(void* hv, const float* q, size_t nq, size_t num_docs, size_t ef_collection,
                          size_t out_stride, uint64_t* labels, float* dists, uint32_t* counts) {
    try {
        RefMv* h = static_cast<RefMv*>(hv);
        for (size_t r = 0; r < nq; r++) {
            hnswlib::MultiVectorSearchStopCondition<unsigned, float> sc(h->space, num_docs, ef_collection);
            auto res = h->alg->searchStopConditionClosest(q + r * h->dim, sc);
            counts[r] = (uint32_t)res.size();
            for (size_t i = 0; i < out_stride; i++) {
                labels[r * out_stride + i] = i < res.size() ? (uint64_t)res[i].second : UINT64_MAX;
                dists[r * out_stride + i] = i < res.size() ? res[i].first : __builtin_inff();
            }
        }
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// BFIndex.knn_query semantics (bindings.cpp:850-929 over bruteforce.h:106-125).
REF_API int ref_bf_search(int space, size_t dim, const float* data, size_t n, const float* q, size_t nq,
                          size_t k, uint64_t* labels, float* dists) {
    try {
        RefSpace s(space, dim);
        hnswlib::BruteforceSearch<float> bf(s.sp, n);
        for (size_t i = 0; i < n; i++) bf.addPoint(data + i * dim, (hnswlib::labeltype)i);
        for (size_t r = 0; r < nq; r++) {
            auto res = bf.searchKnn(q + r * dim, k);
            for (int i = (int)k - 1; i >= 0; i--) {
                labels[r * k + i] = res.top().second;
                dists[r * k + i] = res.top().first;
                res.pop();
            }
        }
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// getRandomLevel sequence for a seed (hnswalg.h:207-211), drawn from a fresh index whose
// generator is seeded as in the constructor (hnswalg.h:117).
REF_API void ref_random_levels(size_t M, size_t seed, size_t count, int* out) {
    hnswlib::L2Space sp(4);
    hnswlib::HierarchicalNSW<float> alg(&sp, 1, M, 200, seed);
    for (size_t i = 0; i < count; i++) out[i] = alg.getRandomLevel(alg.mult_);
}
