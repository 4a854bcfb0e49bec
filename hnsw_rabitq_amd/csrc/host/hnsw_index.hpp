// CPU-side HNSW index: the construction prerequisite of the GPU query path.
//
// The GPU engine searches the graph that the reference's CPU addPoint builds, so this class
// keeps the reference's in-memory layout byte for byte (hnswalg.h:120-143, layout in the
// header comment below) and restates its insertion algorithm:
//   constructor                 hnswalg.h:89-144
//   getRandomLevel              hnswalg.h:207-211
//   searchBaseLayer (build)     hnswalg.h:225-305
//   getNeighborsByHeuristic2    hnswalg.h:443-483
//   mutuallyConnectNewElement   hnswalg.h:506-630 (isUpdate == false branch)
//   addPoint                    hnswalg.h:1152-1266 (new-label branch)
//   saveIndex / loadIndex       hnswalg.h:685-821
// Single-threaded insertion is deterministic and reproduces the reference's file bytes;
// multi-threaded insertion uses the reference's locking scheme and, like the reference, gives
// a valid but schedule-dependent graph.
//
// Level-0 record of element i at level0_ + i * size_data_per_element_:
//   [u16 count][u8 flags (bit0 deleted)][u8 0][u32 nbr[maxM0]] [f32 vec[dim]] [u64 label]
// Upper links of element i: link_lists_[i] + (level-1) * size_links_per_element_,
//   each [u32 count-word][u32 nbr[maxM]].
#pragma once
#include <atomic>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <queue>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "space.hpp"

namespace hnswgpu {

typedef uint32_t tableint;
typedef uint64_t labeltype;

// Arrays handed to the GPU exporter (see gpu_layout.hpp).
struct HostGraphView {
    size_t n = 0, dim = 0, maxM = 0, maxM0 = 0;
    int maxlevel = 0;
    tableint entry = 0;
};

class HnswIndex {
 public:
    // maxM0 = 0: 2 * M, as the reference; another value gives a graph whose level 0 follows the
    // rules of the reference's upper levels when maxM0 = M (used to build those levels)
    HnswIndex(int space, size_t dim, size_t max_elements, size_t M, size_t ef_construction, size_t random_seed,
              size_t maxM0 = 0);
    // Load from a reference-format file (hnswalg.h:715-821).  max_elements == 0 keeps the
    // stored capacity, as the reference does.
    HnswIndex(int space, size_t dim, const std::string& path, size_t max_elements);
    ~HnswIndex();

    HnswIndex(const HnswIndex&) = delete;
    HnswIndex& operator=(const HnswIndex&) = delete;

    // Insert rows; labels may be null (then next_label_ + row).  num_threads <= 0 uses all
    // hardware threads; the first point of an empty index is inserted alone first, and batches
    // of at most 4 * num_threads rows run single-threaded (bindings.cpp:251-304).
    // levels: optional level of each row (else drawn from the index's generator).
    void add_items(const float* data, size_t rows, const labeltype* labels, int num_threads,
                   const int* levels = nullptr);
    // The same, linking only the levels above 0 (the level-0 lists of the new elements stay
    // empty for a level-0 builder to fill).  The upper levels do not depend on level 0, so they
    // are those add_items builds.
    void add_items_upper(const float* data, size_t rows, const labeltype* labels, int num_threads,
                         const int* levels = nullptr);
    // Rows stored with the given levels and no links at any level (empty lists; the entry point
    // and maximum level are left to set_entry), for a builder that fills every level.
    void register_items(const float* data, size_t rows, const labeltype* labels, const int* levels,
                        int num_threads);
    // The levels of the next `rows` insertions, drawn as a one-thread add_items would draw them.
    std::vector<int> draw_levels(size_t rows);
    // Replaces the level-`level` (>= 1) link list of element i.
    void set_links(tableint i, int level, const uint32_t* ids, size_t count);
    void set_entry(tableint entry, int maxlevel);
    // Replaces the level-0 link list of element i (count entries; the delete mark is kept).
    void set_links0(tableint i, const uint32_t* ids, size_t count);
    void save(const std::string& path) const;
    size_t serialized_size() const;
    void serialize(char* out) const;

    void set_ef(size_t ef) { ef_ = ef; }
    size_t ef() const { return ef_; }
    size_t count() const { return cur_element_count_; }
    size_t max_elements() const { return max_elements_; }
    size_t dim() const { return dim_; }
    int space() const { return space_; }
    size_t M() const { return M_; }
    size_t maxM() const { return maxM_; }
    size_t maxM0() const { return maxM0_; }
    size_t ef_construction() const { return ef_construction_; }
    int maxlevel() const { return maxlevel_; }
    tableint entry() const { return enterpoint_node_; }
    int element_level(tableint i) const { return element_levels_[i]; }
    size_t num_deleted() const;
    // markDelete / unmarkDelete (hnswalg.h:852-927): set or clear the delete mark (bit 0 of byte
    // 2 of the level-0 header); the graph is unchanged.
    void mark_deleted(labeltype label);
    void unmark_deleted(labeltype label);

    const uint32_t* linklist0(tableint i) const {
        return (const uint32_t*)(level0_ + i * size_data_per_element_ + offset_level0_);
    }
    const uint32_t* linklist(tableint i, int level) const {
        return (const uint32_t*)(link_lists_[i] + (level - 1) * size_links_per_element_);
    }
    const float* vector(tableint i) const { return (const float*)(level0_ + i * size_data_per_element_ + offset_data_); }
    labeltype label(tableint i) const {
        labeltype l;
        memcpy(&l, level0_ + i * size_data_per_element_ + label_offset_, sizeof(l));
        return l;
    }
    static unsigned short list_count(const uint32_t* p) { return *(const unsigned short*)p; }
    bool is_deleted(tableint i) const { return (((const unsigned char*)linklist0(i))[2] & 0x01) != 0; }

    // Random level sequence check hook (the generator is the reference's
    // std::default_random_engine seeded with random_seed).
    int draw_level() { return random_level(mult_); }

 private:
    typedef std::pair<float, tableint> DistId;
    struct CompareByFirst {
        bool operator()(const DistId& a, const DistId& b) const { return a.first < b.first; }
    };
    typedef std::priority_queue<DistId, std::vector<DistId>, CompareByFirst> MaxHeap;

    struct Visited {
        std::vector<uint16_t> tags;
        uint16_t cur = 0;
        void resize(size_t n) { tags.assign(n, 0); cur = 0; }
        void next() {
            cur++;
            if (cur == 0) {
                std::fill(tags.begin(), tags.end(), 0);
                cur = 1;
            }
        }
    };

    void init_layout();
    int random_level(double reverse_size);
    tableint insert(const float* point, labeltype label, int level = -1);
    MaxHeap search_layer(tableint ep, const float* point, int layer, Visited& vis);
    void select_neighbors(MaxHeap& top, size_t M);
    tableint connect_new_element(const float* point, tableint cur_c, MaxHeap& top, int level);
    Visited& visited_for_thread();

    uint32_t* ll0_mut(tableint i) { return (uint32_t*)(level0_ + i * size_data_per_element_ + offset_level0_); }
    uint32_t* ll_mut(tableint i, int level) {
        return (uint32_t*)(link_lists_[i] + (level - 1) * size_links_per_element_);
    }
    static void set_list_count(uint32_t* p, unsigned short n) { *(unsigned short*)p = n; }
    float dist(const float* a, const float* b) const { return fn_(a, b, dim_); }

    int space_ = 0;
    size_t dim_ = 0;
    DistFn fn_ = nullptr;
    size_t max_elements_ = 0;
    std::atomic<size_t> cur_element_count_{0};
    size_t size_data_per_element_ = 0, size_links_per_element_ = 0, size_links_level0_ = 0;
    size_t M_ = 0, maxM_ = 0, maxM0_ = 0, ef_construction_ = 0, ef_ = 10;
    double mult_ = 0.0;
    int maxlevel_ = -1;
    tableint enterpoint_node_ = (tableint)-1;
    size_t offset_data_ = 0, offset_level0_ = 0, label_offset_ = 0;
    char* level0_ = nullptr;
    std::vector<char*> link_lists_;
    std::vector<int> element_levels_;
    std::vector<std::mutex> link_list_locks_;
    std::mutex global_, label_lookup_lock_, visited_lock_;
    std::unordered_map<labeltype, tableint> label_lookup_;
    std::default_random_engine level_generator_;
    std::vector<Visited*> visited_pool_;
    labeltype next_label_ = 0;
    bool ep_added_ = false;
    bool skip_level0_ = false;
};

}  // namespace hnswgpu
