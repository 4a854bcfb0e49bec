// The searchKnnCloserFirst invariant of the reference's tests/cpp/searchKnnCloserFirst_test.cpp
// (lines 46-68) on the C++ drop-in classes of include/hnswgpu.hpp: for both the brute force and
// the HNSW index, searchKnnCloserFirst returns searchKnn's heap popped from the back.  Also: the
// batched entry point gives the single-query rows, a filter built once gives the functor's
// results, and the brute force equals the k smallest (distance, label) pairs of a scan.
#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdio>
#include <random>
#include <stdexcept>
#include <vector>

#include "../../include/hnswgpu.hpp"

namespace {

int failures = 0;
#define EXPECT(c)                                                            \
    do {                                                                     \
        if (!(c)) {                                                          \
            std::fprintf(stderr, "%s:%d: expectation failed: %s\n", __FILE__, __LINE__, #c); \
            failures++;                                                      \
        }                                                                    \
    } while (0)

struct EvenLabels : hnswgpu::FilterFunctor {
    bool operator()(hnswgpu::labeltype l) override { return l % 2 == 0; }
};

template <class Index>
void closer_first_invariant(const Index& idx, const std::vector<float>& query, size_t nq, size_t d, size_t k,
                            hnswgpu::FilterFunctor* f) {
    for (size_t j = 0; j < nq; ++j) {
        const void* p = query.data() + j * d;
        auto gd = idx.searchKnn(p, k, f);
        auto res = idx.searchKnnCloserFirst(p, k, f);
        EXPECT(gd.size() == res.size());
        size_t t = gd.size();
        while (!gd.empty()) {
            EXPECT(gd.top() == res[--t]);
            gd.pop();
        }
    }
}

void test() {
    const size_t d = 4, n = 100, nq = 10, k = 10;
    std::vector<float> data(n * d), query(nq * d);
    std::mt19937 rng;
    rng.seed(47);
    std::uniform_real_distribution<> distrib;
    for (auto& x : data) x = (float)distrib(rng);
    for (auto& x : query) x = (float)distrib(rng);

    hnswgpu::BruteforceSearch brute(hnswgpu::Space::L2, d, 2 * n);
    hnswgpu::HierarchicalNSW hnsw(hnswgpu::Space::L2, d, 2 * n);
    for (size_t i = 0; i < n; ++i) {
        brute.addPoint(data.data() + d * i, i);
        hnsw.addPoint(data.data() + d * i, i);
    }
    closer_first_invariant(brute, query, nq, d, k, nullptr);
    closer_first_invariant(hnsw, query, nq, d, k, nullptr);
    EvenLabels even;
    closer_first_invariant(brute, query, nq, d, k, &even);
    closer_first_invariant(hnsw, query, nq, d, k, &even);

    // brute force = the k smallest (distance, label) pairs of a scan (bruteforce.h:106-125); the
    // distances here are compared as a set of labels and as sorted values
    for (size_t j = 0; j < nq; ++j) {
        std::vector<std::pair<float, size_t>> all;
        for (size_t i = 0; i < n; ++i) {
            float s = 0;
            for (size_t c = 0; c < d; ++c) {
                const float t = query[j * d + c] - data[i * d + c];
                s += t * t;
            }
            all.emplace_back(s, i);
        }
        std::sort(all.begin(), all.end());
        auto res = brute.searchKnnCloserFirst(query.data() + j * d, k);
        EXPECT(res.size() == k);
        for (size_t i = 0; i < k && i < res.size(); ++i) {
            EXPECT(res[i].second == all[i].second);
            EXPECT(std::abs(res[i].first - all[i].first) <= 1e-6f * std::max(1.0f, all[i].first));
        }
    }

    // one launch for the batch = the single-query rows
    std::vector<uint64_t> bl(nq * k);
    std::vector<float> bd(nq * k);
    hnsw.setEf(40);
    EXPECT(hnsw.searchKnnBatch(query.data(), nq, k, bl.data(), bd.data()));
    for (size_t j = 0; j < nq; ++j) {
        auto res = hnsw.searchKnnCloserFirst(query.data() + j * d, k);
        for (size_t i = 0; i < k; ++i) {
            EXPECT(res[i].second == bl[j * k + i]);
            EXPECT(res[i].first == bd[j * k + i]);
        }
    }
    // a filter built once = the functor evaluated per search
    auto filter = hnsw.makeFilter(even);
    EXPECT(filter->allowed() == n / 2);
    for (size_t j = 0; j < nq; ++j) {
        auto a = hnswgpu::closer_first(hnsw.searchKnn(query.data() + j * d, k, *filter));
        auto b = hnsw.searchKnnCloserFirst(query.data() + j * d, k, &even);
        EXPECT(a == b);
        for (auto& r : a) EXPECT(r.second % 2 == 0);
    }
    // a deleted label is never returned again
    auto first = hnsw.searchKnnCloserFirst(query.data(), k);
    hnsw.markDelete(first[0].second);
    auto after = hnsw.searchKnnCloserFirst(query.data(), k);
    for (auto& r : after) EXPECT(r.second != first[0].second);
    // errors surface as exceptions with the reference's message
    bool threw = false;
    try {
        hnsw.markDelete(first[0].second);
    } catch (const std::runtime_error& e) {
        threw = std::string(e.what()).find("already deleted") != std::string::npos;
    }
    EXPECT(threw);
}

}  // namespace

int main() {
    std::printf("Testing ...\n");
    test();
    if (failures) {
        std::printf("%d expectations failed\n", failures);
        return 1;
    }
    std::printf("Test ok\n");
    return 0;
}
