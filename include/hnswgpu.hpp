// hnswgpu.hpp: C++ drop-in classes for the query path of hnswlib, over the C ABI of hnswgpu.h.
//
// HierarchicalNSW and BruteforceSearch keep the reference's C++ surface for a caller of the query
// path (hnswlib/hnswlib.h:186-221, hnswalg.h, bruteforce.h):
//   searchKnn(query, k, filter)            -> std::priority_queue<std::pair<float, labeltype>>,
//                                             top = farthest (hnswlib.h:192-193)
//   searchKnnCloserFirst(query, k, filter) -> std::vector, closest first (hnswlib.h:203-221)
//   addPoint, setEf, markDelete / unmarkDelete, removePoint, saveIndex, loading from a file
// and add what a GPU needs: searchKnnBatch runs many queries in one launch (a ParallelFor caller,
// bindings.cpp:643-655, would otherwise launch once per row), and a Filter is built once from a
// functor (evaluated once per element) and reused by any number of searches.
//
// No reference header is included: FilterFunctor has the shape of hnswlib::BaseFilterFunctor
// (hnswlib.h:128-132).  Errors throw std::runtime_error with the library's message (the
// reference's texts where the condition is the same).  A search is const and may run
// concurrently with other searches, not with addPoint / markDelete (as in the reference,
// README.md:74,89).  The graph is built on the CPU (addPoint restated); the first search after a
// change exports it to the GPU.
#ifndef HNSWGPU_HPP
#define HNSWGPU_HPP

#include <algorithm>
#include <cstdint>
#include <memory>
#include <mutex>
#include <queue>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "hnswgpu.h"

namespace hnswgpu {

typedef size_t labeltype;

enum class Space { L2 = HG_SPACE_L2, IP = HG_SPACE_IP };

// hnswlib::BaseFilterFunctor (hnswlib.h:128-132): true = the label may be returned.
class FilterFunctor {
 public:
    virtual bool operator()(labeltype) { return true; }
    virtual ~FilterFunctor() {}
};

inline void check(int rc) {
    if (rc != HG_OK) throw std::runtime_error(hg_last_error());
}

typedef std::priority_queue<std::pair<float, labeltype>> ResultHeap;

// Rows of a batch (closest first, padded with label UINT64_MAX) into a heap, top = farthest.
inline ResultHeap row_heap(const uint64_t* lab, const float* dis, size_t k) {
    ResultHeap res;
    for (size_t i = 0; i < k; i++)
        if (lab[i] != UINT64_MAX) res.emplace(dis[i], (labeltype)lab[i]);
    return res;
}

inline std::vector<std::pair<float, labeltype>> closer_first(ResultHeap h) {
    std::vector<std::pair<float, labeltype>> out(h.size());
    size_t n = out.size();
    while (!h.empty()) {
        out[--n] = h.top();
        h.pop();
    }
    return out;
}

class HierarchicalNSW {
 public:
    // A label filter of one export of this index, built once (the functor is called once per
    // element) and reused by searches until the index changes.
    class Filter {
     public:
        ~Filter() {
            if (h_) hg_gpu_filter_free(h_);
        }
        Filter(const Filter&) = delete;
        Filter& operator=(const Filter&) = delete;
        size_t allowed() const {
            size_t n = 0;
            check(hg_gpu_filter_count(h_, &n));
            return n;
        }

     private:
        friend class HierarchicalNSW;
        Filter() = default;
        void* h_ = nullptr;
        const void* owner_gpu_ = nullptr;
    };

    // HierarchicalNSW(space, max_elements, M, ef_construction, random_seed) (hnswalg.h:89-144)
    HierarchicalNSW(Space space, size_t dim, size_t max_elements, size_t M = 16, size_t ef_construction = 200,
                    size_t random_seed = 100, int device = 0)
        : dim_(dim), device_(device) {
        check(hg_hnsw_create((int)space, dim, max_elements, M, ef_construction, random_seed, &cpu_));
    }
    // HierarchicalNSW(space, location, ...) -> loadIndex (hnswalg.h:715-821)
    HierarchicalNSW(Space space, size_t dim, const std::string& location, size_t max_elements = 0, int device = 0)
        : dim_(dim), device_(device) {
        check(hg_hnsw_load((int)space, dim, location.c_str(), max_elements, &cpu_));
    }
    ~HierarchicalNSW() {
        if (gpu_) hg_gpu_index_free(gpu_);
        if (cpu_) hg_hnsw_free(cpu_);
    }
    HierarchicalNSW(const HierarchicalNSW&) = delete;
    HierarchicalNSW& operator=(const HierarchicalNSW&) = delete;

    // addPoint (hnswalg.h:1152-1266), on the CPU; the next search exports the graph again
    void addPoint(const void* datapoint, labeltype label, bool replace_deleted = false) {
        if (replace_deleted) throw std::runtime_error("replacing deleted elements is not supported by this index");
        const uint64_t l = label;
        std::lock_guard<std::mutex> lk(mu_);
        check(hg_hnsw_add_items(cpu_, (const float*)datapoint, 1, &l, 1));
        dirty_ = true;
    }
    // many rows at once (Index.add_items, bindings.cpp:251-304)
    void addPoints(const float* data, size_t rows, const labeltype* labels, int num_threads = 0) {
        std::vector<uint64_t> l(labels, labels + rows);
        std::lock_guard<std::mutex> lk(mu_);
        check(hg_hnsw_add_items(cpu_, data, rows, l.data(), num_threads));
        dirty_ = true;
    }
    void setEf(size_t ef) {  // hnswalg.h:173-175
        std::lock_guard<std::mutex> lk(mu_);
        ef_ = ef;
    }
    size_t getEf() const { return ef_; }
    void markDelete(labeltype label) {  // hnswalg.h:852-890
        std::lock_guard<std::mutex> lk(mu_);
        check(hg_hnsw_mark_deleted(cpu_, label));
        marks_dirty_ = true;
    }
    void unmarkDelete(labeltype label) {
        std::lock_guard<std::mutex> lk(mu_);
        check(hg_hnsw_unmark_deleted(cpu_, label));
        marks_dirty_ = true;
    }
    void saveIndex(const std::string& location) const { check(hg_hnsw_save(cpu_, location.c_str())); }

    std::unique_ptr<Filter> makeFilter(FilterFunctor& f) const {
        void* g = gpu();
        size_t n = 0;
        check(hg_gpu_index_labels(g, nullptr, 0, &n));
        std::vector<uint64_t> all(n), ok;
        check(hg_gpu_index_labels(g, all.data(), n, &n));
        for (uint64_t l : all)
            if (f((labeltype)l)) ok.push_back(l);
        std::unique_ptr<Filter> out(new Filter());
        check(hg_gpu_filter_create(g, ok.data(), ok.size(), &out->h_));
        out->owner_gpu_ = g;
        return out;
    }

    // searchKnn (hnswalg.h:1269-1323): top = farthest; fewer than k results when the index
    // holds fewer reachable elements (the reference returns what it found)
    ResultHeap searchKnn(const void* query, size_t k, FilterFunctor* isIdAllowed = nullptr) const {
        std::vector<uint64_t> lab(k);
        std::vector<float> dis(k);
        if (isIdAllowed) {
            std::unique_ptr<Filter> f = makeFilter(*isIdAllowed);
            search_rows((const float*)query, 1, k, lab.data(), dis.data(), f.get());
        } else {
            search_rows((const float*)query, 1, k, lab.data(), dis.data(), nullptr);
        }
        return row_heap(lab.data(), dis.data(), k);
    }
    ResultHeap searchKnn(const void* query, size_t k, const Filter& filter) const {
        std::vector<uint64_t> lab(k);
        std::vector<float> dis(k);
        search_rows((const float*)query, 1, k, lab.data(), dis.data(), &filter);
        return row_heap(lab.data(), dis.data(), k);
    }
    // searchKnnCloserFirst (hnswlib.h:203-221)
    std::vector<std::pair<float, labeltype>> searchKnnCloserFirst(const void* query, size_t k,
                                                                  FilterFunctor* isIdAllowed = nullptr) const {
        return closer_first(searchKnn(query, k, isIdAllowed));
    }
    // nq queries in one launch: rows of k, closest first, padded with UINT64_MAX / +inf; returns
    // false when some row holds fewer than k results
    bool searchKnnBatch(const float* queries, size_t nq, size_t k, uint64_t* labels, float* dists,
                        const Filter* filter = nullptr) const {
        return search_rows(queries, nq, k, labels, dists, filter);
    }

    size_t dim() const { return dim_; }

 private:
    void* gpu() const {
        std::lock_guard<std::mutex> lk(mu_);
        if (!gpu_ || dirty_) {
            if (gpu_) hg_gpu_index_free(gpu_);
            gpu_ = nullptr;
            check(hg_gpu_index_create(cpu_, device_, &gpu_));
            dirty_ = marks_dirty_ = false;
        } else if (marks_dirty_) {
            check(hg_gpu_refresh_deleted(gpu_, cpu_));
            marks_dirty_ = false;
        }
        return gpu_;
    }
    bool search_rows(const float* q, size_t nq, size_t k, uint64_t* lab, float* dis, const Filter* f) const {
        void* g = gpu();
        if (f && f->owner_gpu_ != g) throw std::runtime_error("the filter was built for another export of the index");
        const int rc = hg_gpu_search_filtered(g, q, nq, k, ef_, f ? f->h_ : nullptr, lab, dis, nullptr);
        if (rc == HG_ERR_SHORT) return false;
        check(rc);
        return true;
    }

    size_t dim_;
    int device_;
    size_t ef_ = 10;
    void* cpu_ = nullptr;
    mutable void* gpu_ = nullptr;
    mutable bool dirty_ = true, marks_dirty_ = false;
    mutable std::mutex mu_;
};

class BruteforceSearch {
 public:
    // BruteforceSearch(space, maxElements) (bruteforce.h:46-56)
    BruteforceSearch(Space space, size_t dim, size_t max_elements, int device = 0) : dim_(dim) {
        check(hg_bf_create((int)space, dim, max_elements, device, &h_));
    }
    BruteforceSearch(Space space, size_t dim, const std::string& location, int device = 0) : dim_(dim) {
        check(hg_bf_load((int)space, dim, location.c_str(), device, &h_));
    }
    ~BruteforceSearch() {
        if (h_) hg_bf_free(h_);
    }
    BruteforceSearch(const BruteforceSearch&) = delete;
    BruteforceSearch& operator=(const BruteforceSearch&) = delete;

    void addPoint(const void* datapoint, labeltype label, bool = false) {  // bruteforce.h:64-83
        const uint64_t l = label;
        check(hg_bf_add(h_, (const float*)datapoint, 1, &l));
    }
    void removePoint(labeltype label) { check(hg_bf_delete(h_, label)); }  // bruteforce.h:85-104
    void saveIndex(const std::string& location) const { check(hg_bf_save(h_, location.c_str())); }
    // HG_PREC_FP32 (default), HG_PREC_BF16 or HG_PREC_FP8: the reduced precisions return the
    // same lists (certified, or answered by the exact path)
    void setPrecision(int precision, size_t rerank_c = 4096) {
        precision_ = precision;
        rerank_c_ = rerank_c;
    }
    size_t count() const {
        size_t n = 0;
        check(hg_bf_count(h_, &n));
        return n;
    }

    // searchKnn (bruteforce.h:106-125): the k smallest (distance, label) pairs, top = farthest
    ResultHeap searchKnn(const void* query, size_t k, FilterFunctor* isIdAllowed = nullptr) const {
        const size_t kk = std::min(k, count());
        std::vector<uint64_t> lab(kk);
        std::vector<float> dis(kk);
        if (kk == 0) return ResultHeap();
        if (isIdAllowed) {
            std::vector<uint64_t> all(count()), ok;
            check(hg_bf_labels(h_, all.data(), all.size()));
            for (uint64_t l : all)
                if ((*isIdAllowed)((labeltype)l)) ok.push_back(l);
            void* f = nullptr;
            check(hg_bf_filter_create(h_, ok.data(), ok.size(), &f));
            const int rc = hg_bf_search_filtered(h_, (const float*)query, 1, kk, f, lab.data(), dis.data());
            hg_bf_filter_free(f);
            if (rc != HG_ERR_SHORT) check(rc);
        } else {
            searchKnnBatch((const float*)query, 1, kk, lab.data(), dis.data());
        }
        return row_heap(lab.data(), dis.data(), kk);
    }
    std::vector<std::pair<float, labeltype>> searchKnnCloserFirst(const void* query, size_t k,
                                                                  FilterFunctor* isIdAllowed = nullptr) const {
        return closer_first(searchKnn(query, k, isIdAllowed));
    }
    // nq queries in one launch (k <= count())
    void searchKnnBatch(const float* queries, size_t nq, size_t k, uint64_t* labels, float* dists) const {
        check(hg_bf_search_precision(h_, queries, nq, k, precision_, rerank_c_, labels, dists));
    }

 private:
    size_t dim_;
    void* h_ = nullptr;
    int precision_ = HG_PREC_FP32;
    size_t rerank_c_ = 4096;
};

}  // namespace hnswgpu

#endif
