// Host-side code under AddressSanitizer and UndefinedBehaviorSanitizer: the C++ graph builder and
// its on-disk format (hnsw_rabitq_amd/csrc/host) and the CPU oracle (oracle/oracle.cpp), driven
// through construction (one and several threads, upper levels only, registered rows), save /
// load / serialize round trips, delete marks, every oracle search mode, brute force, the FP64
// pairs and the network's forward and gradient, plus the error paths: truncated and garbled
// index files and insertion past the capacity.  Built and run by tests/test_host_sanitize.py with
// g++ -fsanitize=address,undefined; any sanitizer report or failed check exits non-zero.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "host/hnsw_index.hpp"
#include "oracle.h"

using hnswgpu::HnswIndex;

static int failures = 0;
#define CHECK(c)                                                       \
    do {                                                               \
        if (!(c)) {                                                    \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            failures++;                                                \
        }                                                              \
    } while (0)

static std::vector<float> random_rows(size_t n, size_t dim, unsigned seed, bool integers) {
    std::mt19937 g(seed);
    std::uniform_real_distribution<float> u(0.f, 1.f);
    std::vector<float> v(n * dim);
    for (auto& x : v) x = integers ? std::floor(u(g) * 40.f) : u(g);
    return v;
}

static std::vector<char> bytes_of(const HnswIndex& h) {
    std::vector<char> b(h.serialized_size());
    h.serialize(b.data());
    return b;
}

static void searches(const std::vector<char>& blob, int space, size_t dim, const std::vector<float>& q, size_t nq,
                     size_t n) {
    void* o = or_hnsw_load_bytes(blob.data(), blob.size());
    CHECK(o != nullptr);
    if (!o) return;
    const size_t k = 10, kmax = 16;
    std::vector<uint64_t> lab(nq * kmax);
    std::vector<float> dis(nq * kmax);
    std::vector<int64_t> ctr(nq * 4);
    std::vector<uint8_t> ties(nq);
    std::vector<int32_t> efq(nq), kq(nq), caps(nq);
    for (size_t i = 0; i < nq; i++) {
        efq[i] = 1 + (int32_t)(i * 37 % 200);
        kq[i] = 1 + (int32_t)(i % kmax);
        caps[i] = (int32_t)(i % 50) - 5;
    }
    for (int mode = 0; mode < 2; mode++) {
        CHECK(or_hnsw_search(o, space, q.data(), nq, k, 50, nullptr, nullptr, k, mode, 2, lab.data(), dis.data(),
                             ctr.data(), ties.data()) >= 0);
        CHECK(or_hnsw_search(o, space, q.data(), nq, kmax, 200, efq.data(), kq.data(), kmax, mode, 3, lab.data(),
                             dis.data(), ctr.data(), nullptr) >= 0);
        CHECK(or_hnsw_search_hopcap(o, space, q.data(), nq, k, 100, caps.data(), k, mode, 2, lab.data(), dis.data(),
                                    ctr.data(), nullptr) >= 0);
    }
    std::vector<uint64_t> labels_by_id(n);
    std::vector<uint8_t> deleted(n), allowed(n);
    or_hnsw_labels(o, labels_by_id.data(), deleted.data());
    for (size_t i = 0; i < n; i++) allowed[i] = (i % 3) == 0;
    CHECK(or_hnsw_search_allowed(o, space, q.data(), nq, k, 60, nullptr, nullptr, k, 1, 2, allowed.data(), lab.data(),
                                 dis.data(), ctr.data(), nullptr) >= 0);
    std::vector<uint32_t> counts(nq), docs(n);
    for (size_t i = 0; i < n; i++) docs[i] = (uint32_t)(i % 97);
    CHECK(or_hnsw_search_epsilon(o, space, q.data(), nq, space == 0 ? 0.5f : 0.9f, 5, 40, 0, 2, allowed.data(), kmax,
                                 lab.data(), dis.data(), counts.data(), ctr.data(), ties.data()) >= 0);
    CHECK(or_hnsw_search_multivector(o, space, q.data(), nq, docs.data(), 8, 20, 1, 2, nullptr, kmax, lab.data(),
                                     dis.data(), counts.data(), ctr.data(), nullptr) >= 0);
    or_hnsw_free(o);
}

int main() {
    const std::string dir = std::getenv("SANITIZE_TMP") ? std::getenv("SANITIZE_TMP") : "/tmp";
    // construction, serialisation, reload, searches
    for (int space = 0; space < 2; space++) {
        for (size_t dim : {20u, 32u}) {
            const size_t n = 2500, nq = 40;
            const auto data = random_rows(n, dim, 7 + (unsigned)dim, space == 0 && dim == 32);
            const auto q = random_rows(nq, dim, 99, false);
            for (int threads : {1, 4}) {
                HnswIndex h(space, dim, n, 8, 60, 100);
                h.add_items(data.data(), n / 2, nullptr, threads);
                h.add_items(data.data() + n / 2 * dim, n - n / 2, nullptr, threads);
                const auto blob = bytes_of(h);
                const std::string path = dir + "/sanitize_index.bin";
                h.save(path);
                HnswIndex r(space, dim, path, 0);
                CHECK(bytes_of(r) == blob);
                HnswIndex bigger(space, dim, path, n + 100);
                CHECK(bigger.max_elements() == n + 100);
                std::remove(path.c_str());
                h.mark_deleted(3);
                h.mark_deleted(17);
                CHECK(h.num_deleted() == 2);
                h.unmark_deleted(17);
                CHECK(h.num_deleted() == 1);
                searches(bytes_of(h), space, dim, q, nq, n);
                // past the capacity: the insertion fails, the index stays usable
                bool threw = false;
                try {
                    h.add_items(data.data(), 1, nullptr, 1);
                } catch (const std::exception&) {
                    threw = true;
                }
                CHECK(threw);
                CHECK(h.count() == n);
            }
        }
    }
    // the GPU builder's host steps: upper levels only, registered rows, replaced link lists
    {
        const size_t n = 1200, dim = 16;
        const auto data = random_rows(n, dim, 5, false);
        HnswIndex h(0, dim, n, 8, 40, 100);
        const auto levels = h.draw_levels(n);
        h.add_items_upper(data.data(), n, nullptr, 2, levels.data());
        std::vector<uint32_t> row;
        for (size_t i = 0; i < n; i++) {
            row.clear();
            for (size_t j = 1; j <= 4; j++) row.push_back((uint32_t)((i + j * 97) % n));
            h.set_links0((uint32_t)i, row.data(), row.size());
        }
        searches(bytes_of(h), 0, dim, random_rows(20, dim, 6, false), 20, n);
        HnswIndex g(0, dim, n, 8, 40, 100);
        g.register_items(data.data(), n, nullptr, levels.data(), 2);
        g.set_entry(0, levels[0]);
        CHECK(g.count() == n);
    }
    // corrupted files: every truncation of a small index, and garbled headers, must be refused
    {
        const size_t n = 60, dim = 8;
        const auto data = random_rows(n, dim, 11, false);
        HnswIndex h(0, dim, n, 4, 20, 100);
        h.add_items(data.data(), n, nullptr, 1);
        const auto blob = bytes_of(h);
        const std::string path = dir + "/sanitize_bad.bin";
        int refused = 0, tried = 0;
        for (size_t cut = 0; cut < blob.size(); cut += 7) {
            std::ofstream(path, std::ios::binary).write(blob.data(), (std::streamsize)cut);
            tried++;
            try {
                HnswIndex r(0, dim, path, 0);
            } catch (const std::exception&) {
                refused++;
            }
            void* o = or_hnsw_load_bytes(blob.data(), cut);
            if (o) or_hnsw_free(o);
        }
        CHECK(refused == tried);
        std::mt19937 g(3);
        for (int t = 0; t < 200; t++) {
            std::vector<char> b = blob;
            for (int j = 0; j < 4; j++) b[g() % 96] = (char)g();
            std::ofstream(path, std::ios::binary).write(b.data(), (std::streamsize)b.size());
            try {
                HnswIndex r(0, dim, path, 0);
            } catch (const std::exception&) {
            }
            void* o = or_hnsw_load_bytes(b.data(), b.size());
            if (o) or_hnsw_free(o);
        }
        std::remove(path.c_str());
    }
    // brute force, FP64 pairs, the network
    {
        const size_t n = 700, dim = 24, nq = 9, k = 13;
        const auto data = random_rows(n, dim, 21, false);
        const auto q = random_rows(nq, dim, 22, false);
        std::vector<uint64_t> lab(nq * k);
        std::vector<float> dis(nq * k);
        for (int space = 0; space < 2; space++)
            CHECK(or_bf_search(space, dim, data.data(), nullptr, n, q.data(), nq, k, 2, lab.data(), dis.data()) >= 0);
        std::vector<int64_t> qi = {0, 3, 8}, ri = {1, 600, 699};
        std::vector<double> out(3);
        or_fp64_pairs(0, dim, q.data(), data.data(), qi.data(), ri.data(), 3, out.data());
        CHECK(out[0] >= 0.0);
        const int in = 12, hid = 16;
        const int np = or_mlp_param_count(in, hid);
        std::vector<float> params(np), x(5 * in), y = {0, 1, 1, 0, 1}, logits(5);
        std::vector<double> grad(np);
        std::mt19937 g(4);
        std::normal_distribution<float> nd(0.f, 0.3f);
        for (auto& p : params) p = nd(g);
        for (auto& v : x) v = nd(g);
        or_mlp_forward(params.data(), in, hid, x.data(), 5, logits.data());
        const double loss = or_mlp_loss_grad(params.data(), in, hid, x.data(), y.data(), 5, grad.data());
        CHECK(std::isfinite(loss));
    }
    std::printf("host sanitize: %s\n", failures ? "FAILED" : "ok");
    return failures ? 1 : 0;
}
