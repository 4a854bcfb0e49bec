// CPU construction and persistence of the HNSW graph the GPU engine searches.
// See hnsw_index.hpp for the map to the reference routines this restates.
#include "hnsw_index.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <exception>
#include <fstream>
#include <iostream>
#include <thread>

namespace hnswgpu {

namespace {

template <typename T>
void put(std::vector<char>& out, const T& v) {
    const char* p = (const char*)&v;
    out.insert(out.end(), p, p + sizeof(T));
}

template <typename T>
void put_raw(char*& out, const T& v) {
    memcpy(out, &v, sizeof(T));
    out += sizeof(T);
}

// ParallelFor over [start, end) with an atomic work counter; the first exception stops the
// batch and is rethrown (bindings.cpp:22-74).
template <class F>
void parallel_for(size_t start, size_t end, int num_threads, F fn) {
    if (num_threads <= 0) num_threads = (int)std::thread::hardware_concurrency();
    if (num_threads == 1) {
        for (size_t i = start; i < end; i++) fn(i);
        return;
    }
    std::atomic<size_t> current(start);
    std::exception_ptr last;
    std::mutex last_mu;
    std::vector<std::thread> threads;
    for (int t = 0; t < num_threads; t++) {
        threads.emplace_back([&] {
            for (;;) {
                size_t id = current.fetch_add(1);
                if (id >= end) break;
                try {
                    fn(id);
                } catch (...) {
                    std::lock_guard<std::mutex> g(last_mu);
                    last = std::current_exception();
                    current = end;
                    break;
                }
            }
        });
    }
    for (auto& th : threads) th.join();
    if (last) std::rethrow_exception(last);
}

}  // namespace

HnswIndex::HnswIndex(int space, size_t dim, size_t max_elements, size_t M, size_t ef_construction, size_t random_seed,
                     size_t maxM0)
    : space_(space),
      dim_(dim),
      fn_(distance_fn(space)),
      max_elements_(max_elements),
      element_levels_(max_elements),
      link_list_locks_(max_elements) {
    if (M <= 10000) {
        M_ = M;
    } else {
        std::cerr << "warning: M parameter exceeds 10000 which may lead to adverse effects." << std::endl;
        std::cerr << "         Cap to 10000 will be applied for the rest of the processing." << std::endl;
        M_ = 10000;
    }
    maxM_ = M_;
    maxM0_ = maxM0 ? maxM0 : M_ * 2;
    ef_construction_ = std::max(ef_construction, M_);
    ef_ = 10;
    level_generator_.seed(random_seed);
    init_layout();
    level0_ = (char*)malloc(std::max<size_t>(max_elements_, 1) * size_data_per_element_);
    if (!level0_) throw std::runtime_error("Not enough memory");
    link_lists_.assign(max_elements_, nullptr);
    mult_ = 1 / log(1.0 * M_);
}

void HnswIndex::init_layout() {
    size_links_level0_ = maxM0_ * sizeof(tableint) + sizeof(uint32_t);
    size_data_per_element_ = size_links_level0_ + dim_ * sizeof(float) + sizeof(labeltype);
    offset_data_ = size_links_level0_;
    label_offset_ = size_links_level0_ + dim_ * sizeof(float);
    offset_level0_ = 0;
    size_links_per_element_ = maxM_ * sizeof(tableint) + sizeof(uint32_t);
}

HnswIndex::HnswIndex(int space, size_t dim, const std::string& path, size_t max_elements_i)
    : space_(space), dim_(dim), fn_(distance_fn(space)) {
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) throw std::runtime_error("Cannot open file");
    in.seekg(0, in.end);
    size_t total = (size_t)in.tellg();
    in.seekg(0, in.beg);
    std::vector<char> buf(total);
    in.read(buf.data(), (std::streamsize)total);
    const char* p = buf.data();
    const char* end = p + total;
    auto rd = [&](void* dst, size_t n) {
        if ((size_t)(end - p) < n) throw std::runtime_error("Index seems to be corrupted or unsupported");
        memcpy(dst, p, n);
        p += n;
    };
    size_t offset_level0, max_elements, cur_count, size_per, label_off, off_data, maxM, maxM0, M, efc;
    int maxlevel;
    tableint entry;
    double mult;
    rd(&offset_level0, 8);
    rd(&max_elements, 8);
    rd(&cur_count, 8);
    rd(&size_per, 8);
    rd(&label_off, 8);
    rd(&off_data, 8);
    rd(&maxlevel, 4);
    rd(&entry, 4);
    rd(&maxM, 8);
    rd(&maxM0, 8);
    rd(&M, 8);
    rd(&mult, 8);
    rd(&efc, 8);
    // Whole-file size check as in loadIndex (hnswalg.h:753-769), with the products checked
    // for overflow (a garbled count must not wrap past the check).
    {
        const char* q = p;
        if (size_per == 0 || cur_count > (size_t)(end - q) / size_per)
            throw std::runtime_error("Index seems to be corrupted or unsupported");
        q += cur_count * size_per;
        for (size_t i = 0; i < cur_count; i++) {
            if (q >= end) throw std::runtime_error("Index seems to be corrupted or unsupported");
            uint32_t sz;
            if ((size_t)(end - q) < 4) throw std::runtime_error("Index seems to be corrupted or unsupported");
            memcpy(&sz, q, 4);
            q += 4;
            if ((size_t)(end - q) < sz) throw std::runtime_error("Index seems to be corrupted or unsupported");
            q += sz;
        }
        if (q != end) throw std::runtime_error("Index seems to be corrupted or unsupported");
    }
    if (label_off - off_data != dim * sizeof(float) || size_per != label_off + sizeof(labeltype))
        throw std::runtime_error("Wrong dimensionality of the vectors");
    max_elements_ = max_elements_i < cur_count ? max_elements : max_elements_i;
    // a stored capacity below the stored count, or an entry point or level that the lists do not
    // hold, is a damaged file (the reference would index past its arrays)
    if (max_elements_ < cur_count || (cur_count > 0 && entry >= cur_count) || maxlevel < 0 || maxM0 == 0 || maxM == 0)
        throw std::runtime_error("Index seems to be corrupted or unsupported");
    M_ = M;
    maxM_ = maxM;
    maxM0_ = maxM0;
    ef_construction_ = efc;
    mult_ = mult;
    maxlevel_ = maxlevel;
    enterpoint_node_ = entry;
    ef_ = 10;
    init_layout();
    if (size_per != size_data_per_element_ || offset_level0 != 0)
        throw std::runtime_error("Index seems to be corrupted or unsupported");
    if (max_elements_ > ((size_t)1 << 40) / size_data_per_element_)
        throw std::runtime_error("Index seems to be corrupted or unsupported");  // capacity beyond any host
    level0_ = (char*)malloc(std::max<size_t>(max_elements_, 1) * size_data_per_element_);
    if (!level0_) throw std::runtime_error("out of host memory for the level-0 records of the loaded index");
    // a constructor that throws runs no destructor: what was allocated is freed here
    try {
        rd(level0_, cur_count * size_data_per_element_);
        link_lists_.assign(max_elements_, nullptr);
        element_levels_.assign(max_elements_, 0);
        std::vector<std::mutex>(max_elements_).swap(link_list_locks_);
        for (size_t i = 0; i < cur_count; i++) {
            label_lookup_[label((tableint)i)] = (tableint)i;
            uint32_t sz;
            rd(&sz, 4);
            if (sz == 0) {
                element_levels_[i] = 0;
            } else {
                if (sz % size_links_per_element_ != 0) throw std::runtime_error("Index seems to be corrupted or unsupported");
                element_levels_[i] = (int)(sz / size_links_per_element_);
                link_lists_[i] = (char*)malloc(sz);
                if (!link_lists_[i]) throw std::runtime_error("out of host memory for an upper-level link list of the loaded index");
                rd(link_lists_[i], sz);
            }
        }
        // every list within its capacity, every id an element of the index
        auto bad = [&]() { return std::runtime_error("Index seems to be corrupted or unsupported"); };
        for (size_t i = 0; i < cur_count; i++) {
            const uint32_t* l0 = linklist0((tableint)i);
            const size_t c0 = *(const uint16_t*)l0;
            if (c0 > maxM0_) throw bad();
            for (size_t j = 0; j < c0; j++)
                if (l0[1 + j] >= cur_count) throw bad();
            for (int lv = 1; lv <= element_levels_[i]; lv++) {
                const uint32_t* l = linklist((tableint)i, lv);
                const size_t c = *(const uint16_t*)l;
                if (c > maxM_) throw bad();
                for (size_t j = 0; j < c; j++)
                    if (l[1 + j] >= cur_count) throw bad();
            }
        }
        if (cur_count > 0 && element_levels_[entry] < maxlevel) throw bad();
    } catch (...) {
        for (char* l : link_lists_) free(l);
        link_lists_.clear();
        free(level0_);
        level0_ = nullptr;
        throw;
    }
    cur_element_count_ = cur_count;
    next_label_ = cur_count;
    ep_added_ = true;
}

HnswIndex::~HnswIndex() {
    for (size_t i = 0; i < link_lists_.size(); i++) free(link_lists_[i]);
    free(level0_);
    for (auto* v : visited_pool_) delete v;
}

size_t HnswIndex::num_deleted() const {
    size_t n = 0;
    for (size_t i = 0; i < cur_element_count_; i++) n += is_deleted((tableint)i) ? 1 : 0;
    return n;
}

void HnswIndex::mark_deleted(labeltype label) {
    tableint id;
    {
        std::lock_guard<std::mutex> g(label_lookup_lock_);
        auto it = label_lookup_.find(label);
        if (it == label_lookup_.end()) throw std::runtime_error("Label not found");
        id = it->second;
    }
    unsigned char* flags = (unsigned char*)ll0_mut(id) + 2;
    if (*flags & 0x01) throw std::runtime_error("The requested to delete element is already deleted");
    *flags |= 0x01;
}

void HnswIndex::unmark_deleted(labeltype label) {
    tableint id;
    {
        std::lock_guard<std::mutex> g(label_lookup_lock_);
        auto it = label_lookup_.find(label);
        if (it == label_lookup_.end()) throw std::runtime_error("Label not found");
        id = it->second;
    }
    unsigned char* flags = (unsigned char*)ll0_mut(id) + 2;
    if (!(*flags & 0x01)) throw std::runtime_error("The requested to undelete element is not deleted");
    *flags &= (unsigned char)~0x01;
}

int HnswIndex::random_level(double reverse_size) {
    std::uniform_real_distribution<double> distribution(0.0, 1.0);
    double r = -log(distribution(level_generator_)) * reverse_size;
    return (int)r;
}

HnswIndex::Visited& HnswIndex::visited_for_thread() {
    std::lock_guard<std::mutex> g(visited_lock_);
    Visited* v;
    if (visited_pool_.empty()) {
        v = new Visited();
        v->resize(max_elements_);
    } else {
        v = visited_pool_.back();
        visited_pool_.pop_back();
    }
    return *v;
}

HnswIndex::MaxHeap HnswIndex::search_layer(tableint ep, const float* point, int layer, Visited& vis) {
    vis.next();
    const uint16_t tag = vis.cur;
    MaxHeap top, cand;
    float d0 = dist(point, vector(ep));
    top.emplace(d0, ep);
    float lower_bound = d0;
    cand.emplace(-d0, ep);
    vis.tags[ep] = tag;
    while (!cand.empty()) {
        DistId cp = cand.top();
        if ((-cp.first) > lower_bound && top.size() == ef_construction_) break;
        cand.pop();
        tableint cur = cp.second;
        std::unique_lock<std::mutex> lock(link_list_locks_[cur]);
        const uint32_t* data = layer == 0 ? linklist0(cur) : linklist(cur, layer);
        size_t size = list_count(data);
        const tableint* datal = data + 1;
        for (size_t j = 0; j < size; j++) {
            tableint cid = datal[j];
            if (vis.tags[cid] == tag) continue;
            vis.tags[cid] = tag;
            float d1 = dist(point, vector(cid));
            if (top.size() < ef_construction_ || lower_bound > d1) {
                cand.emplace(-d1, cid);
                if (!is_deleted(cid)) top.emplace(d1, cid);
                if (top.size() > ef_construction_) top.pop();
                if (!top.empty()) lower_bound = top.top().first;
            }
        }
    }
    return top;
}

void HnswIndex::select_neighbors(MaxHeap& top, size_t M) {
    if (top.size() < M) return;
    // closest first; pairs compare (−dist, id) lexicographically as in hnswalg.h:449
    std::priority_queue<std::pair<float, tableint>> closest;
    std::vector<std::pair<float, tableint>> keep;
    while (!top.empty()) {
        closest.emplace(-top.top().first, top.top().second);
        top.pop();
    }
    while (!closest.empty()) {
        if (keep.size() >= M) break;
        std::pair<float, tableint> cp = closest.top();
        float dq = -cp.first;
        closest.pop();
        bool good = true;
        for (const auto& sp : keep) {
            float cd = dist(vector(sp.second), vector(cp.second));
            if (cd < dq) {
                good = false;
                break;
            }
        }
        if (good) keep.push_back(cp);
    }
    for (const auto& cp : keep) top.emplace(-cp.first, cp.second);
}

tableint HnswIndex::connect_new_element(const float* point, tableint cur_c, MaxHeap& top, int level) {
    (void)point;
    size_t mcurmax = level ? maxM_ : maxM0_;
    select_neighbors(top, M_);
    if (top.size() > M_) throw std::runtime_error("internal: the neighbour heuristic kept more than M links");
    std::vector<tableint> selected;
    selected.reserve(M_);
    while (!top.empty()) {
        selected.push_back(top.top().second);
        top.pop();
    }
    tableint next_closest = selected.back();
    {
        uint32_t* ll_cur = level == 0 ? ll0_mut(cur_c) : ll_mut(cur_c, level);
        if (*ll_cur) throw std::runtime_error("internal: a new element's link list is not empty");
        set_list_count(ll_cur, (unsigned short)selected.size());
        tableint* data = ll_cur + 1;
        for (size_t i = 0; i < selected.size(); i++) {
            if (data[i]) throw std::runtime_error("internal: a new element's link slot is not empty");
            if (level > element_levels_[selected[i]])
                throw std::runtime_error("internal: a link on a level above the element's own");
            data[i] = selected[i];
        }
    }
    for (size_t i = 0; i < selected.size(); i++) {
        tableint other = selected[i];
        std::unique_lock<std::mutex> lock(link_list_locks_[other]);
        uint32_t* ll_other = level == 0 ? ll0_mut(other) : ll_mut(other, level);
        size_t sz = list_count(ll_other);
        if (sz > mcurmax) throw std::runtime_error("internal: a link list holds more entries than its capacity");
        if (other == cur_c) throw std::runtime_error("internal: an element selected itself as a neighbour");
        if (level > element_levels_[other]) throw std::runtime_error("internal: a link on a level above the element's own");
        tableint* data = ll_other + 1;
        if (sz < mcurmax) {
            data[sz] = cur_c;
            set_list_count(ll_other, (unsigned short)(sz + 1));
        } else {
            float d_max = dist(vector(cur_c), vector(other));
            MaxHeap cands;
            cands.emplace(d_max, cur_c);
            for (size_t j = 0; j < sz; j++) cands.emplace(dist(vector(data[j]), vector(other)), data[j]);
            select_neighbors(cands, mcurmax);
            int indx = 0;
            while (!cands.empty()) {
                data[indx] = cands.top().second;
                cands.pop();
                indx++;
            }
            set_list_count(ll_other, (unsigned short)indx);
        }
    }
    return next_closest;
}

tableint HnswIndex::insert(const float* point, labeltype lab, int level) {
    tableint cur_c;
    {
        std::lock_guard<std::mutex> g(label_lookup_lock_);
        if (label_lookup_.count(lab))
            throw std::runtime_error("Replacing the vector of an existing label is not supported by this index");
        if (cur_element_count_ >= max_elements_)
            throw std::runtime_error("The number of elements exceeds the specified limit");
        cur_c = (tableint)cur_element_count_;
        cur_element_count_++;
        label_lookup_[lab] = cur_c;
    }
    std::unique_lock<std::mutex> lock_el(link_list_locks_[cur_c]);
    int curlevel = level;
    if (curlevel < 0) {
        std::lock_guard<std::mutex> g(visited_lock_);  // the generator is shared by all inserters
        curlevel = random_level(mult_);
    }
    element_levels_[cur_c] = curlevel;

    // the element's own record (private to this inserter: its lock is held)
    memset(level0_ + cur_c * size_data_per_element_ + offset_level0_, 0, size_data_per_element_);
    memcpy(level0_ + cur_c * size_data_per_element_ + label_offset_, &lab, sizeof(labeltype));
    memcpy(level0_ + cur_c * size_data_per_element_ + offset_data_, point, dim_ * sizeof(float));
    if (curlevel) {
        link_lists_[cur_c] = (char*)malloc(size_links_per_element_ * curlevel + 1);
        if (!link_lists_[cur_c]) throw std::runtime_error("out of host memory for an upper-level link list");
        memset(link_lists_[cur_c], 0, size_links_per_element_ * curlevel + 1);
    }

    std::unique_lock<std::mutex> templock(global_);
    int maxlevelcopy = maxlevel_;
    // the entry point is read under the lock that its writers hold (an inserter that raises the
    // top level keeps the lock to the end and writes it there)
    tableint curr_obj = enterpoint_node_;
    if (curlevel <= maxlevelcopy) templock.unlock();

    if ((int32_t)curr_obj != -1 && !(skip_level0_ && curlevel == 0)) {
        if (curlevel < maxlevelcopy) {
            float curdist = dist(point, vector(curr_obj));
            for (int level = maxlevelcopy; level > curlevel; level--) {
                bool changed = true;
                while (changed) {
                    changed = false;
                    std::unique_lock<std::mutex> lock(link_list_locks_[curr_obj]);
                    const uint32_t* data = linklist(curr_obj, level);
                    int size = list_count(data);
                    const tableint* datal = data + 1;
                    for (int i = 0; i < size; i++) {
                        tableint cand = datal[i];
                        if (cand > max_elements_) throw std::runtime_error("cand error");
                        float d = dist(point, vector(cand));
                        if (d < curdist) {
                            curdist = d;
                            curr_obj = cand;
                            changed = true;
                        }
                    }
                }
            }
        }
        Visited& vis = visited_for_thread();
        try {
            for (int level = std::min(curlevel, maxlevelcopy); level >= (skip_level0_ ? 1 : 0); level--) {
                MaxHeap top = search_layer(curr_obj, point, level, vis);
                curr_obj = connect_new_element(point, cur_c, top, level);
            }
        } catch (...) {
            std::lock_guard<std::mutex> g(visited_lock_);
            visited_pool_.push_back(&vis);
            throw;
        }
        std::lock_guard<std::mutex> g(visited_lock_);
        visited_pool_.push_back(&vis);
    } else if ((int32_t)curr_obj == -1) {
        enterpoint_node_ = 0;
        maxlevel_ = curlevel;
    }
    if (curlevel > maxlevelcopy) {
        enterpoint_node_ = cur_c;
        maxlevel_ = curlevel;
    }
    return cur_c;
}

void HnswIndex::add_items(const float* data, size_t rows, const labeltype* labels, int num_threads,
                          const int* levels) {
    if (num_threads <= 0) num_threads = (int)std::thread::hardware_concurrency();
    if (rows <= (size_t)num_threads * 4) num_threads = 1;
    size_t start = 0;
    labeltype base = next_label_;
    if (!ep_added_ && rows > 0) {
        insert(data, labels ? labels[0] : base, levels ? levels[0] : -1);
        start = 1;
        ep_added_ = true;
    }
    parallel_for(start, rows, num_threads, [&](size_t row) {
        insert(data + row * dim_, labels ? labels[row] : base + row, levels ? levels[row] : -1);
    });
    next_label_ += rows;
}

void HnswIndex::add_items_upper(const float* data, size_t rows, const labeltype* labels, int num_threads,
                                const int* levels) {
    skip_level0_ = true;
    try {
        add_items(data, rows, labels, num_threads, levels);
    } catch (...) {
        skip_level0_ = false;
        throw;
    }
    skip_level0_ = false;
}

void HnswIndex::register_items(const float* data, size_t rows, const labeltype* labels, const int* levels,
                               int num_threads) {
    if (num_threads <= 0) num_threads = (int)std::thread::hardware_concurrency();
    const size_t s0 = cur_element_count_;
    if (rows > max_elements_ - std::min(s0, max_elements_))
        throw std::runtime_error("The number of elements exceeds the specified limit");
    const labeltype base = next_label_;
    // labels first, one thread (the table is reserved once): a label already present throws
    // before anything is stored
    label_lookup_.reserve(s0 + rows);
    for (size_t r = 0; r < rows; r++) {
        const labeltype lab = labels ? labels[r] : base + r;
        if (!label_lookup_.emplace(lab, (tableint)(s0 + r)).second) {
            for (size_t q = 0; q < r; q++) label_lookup_.erase(labels ? labels[q] : base + q);
            throw std::runtime_error("Replacing the vector of an existing label is not supported by this index");
        }
    }
    // element s0 + r is row r: its record, level and empty lists (no other thread reads them),
    // in chunks of 4,096 rows
    const size_t chunk = 4096;
    parallel_for(0, (rows + chunk - 1) / chunk, num_threads, [&](size_t ch) {
        for (size_t r = ch * chunk; r < std::min(rows, (ch + 1) * chunk); r++) {
            const tableint c = (tableint)(s0 + r);
            const labeltype lab = labels ? labels[r] : base + r;
            const int l = levels[r];
            element_levels_[c] = l;
            char* rec = level0_ + c * size_data_per_element_;
            memset(rec + offset_level0_, 0, size_data_per_element_);
            memcpy(rec + label_offset_, &lab, sizeof(labeltype));
            memcpy(rec + offset_data_, data + r * dim_, dim_ * sizeof(float));
            if (l) {
                link_lists_[c] = (char*)malloc(size_links_per_element_ * l + 1);
                if (!link_lists_[c]) throw std::runtime_error("out of host memory for an upper-level link list");
                memset(link_lists_[c], 0, size_links_per_element_ * l + 1);
            }
        }
    });
    cur_element_count_ = s0 + rows;
    next_label_ += rows;
    if (rows > 0) ep_added_ = true;
}

std::vector<int> HnswIndex::draw_levels(size_t rows) {
    std::vector<int> lv(rows);
    std::lock_guard<std::mutex> g(visited_lock_);
    for (auto& l : lv) l = random_level(mult_);
    return lv;
}

void HnswIndex::set_links(tableint i, int level, const uint32_t* ids, size_t count) {
    if (i >= cur_element_count_ || level < 1 || level > element_levels_[i] || count > maxM_)
        throw std::runtime_error("internal: a link list holds more entries than its capacity");
    uint32_t* ll = ll_mut(i, level);
    set_list_count(ll, (unsigned short)count);
    for (size_t j = 0; j < maxM_; j++) ll[1 + j] = j < count ? ids[j] : 0u;
}

void HnswIndex::set_entry(tableint entry, int maxlevel) {
    std::lock_guard<std::mutex> g(global_);
    enterpoint_node_ = entry;
    maxlevel_ = maxlevel;
}

void HnswIndex::set_links0(tableint i, const uint32_t* ids, size_t count) {
    if (i >= cur_element_count_ || count > maxM0_) throw std::runtime_error("internal: a link list holds more entries than its capacity");
    uint32_t* ll = ll0_mut(i);
    set_list_count(ll, (unsigned short)count);
    for (size_t j = 0; j < maxM0_; j++) ll[1 + j] = j < count ? ids[j] : 0u;
}

size_t HnswIndex::serialized_size() const {
    size_t size = 8 * 6 + 4 + 4 + 8 * 3 + 8 + 8;
    size += cur_element_count_ * size_data_per_element_;
    for (size_t i = 0; i < cur_element_count_; i++) {
        size += 4;
        if (element_levels_[i] > 0) size += size_links_per_element_ * element_levels_[i];
    }
    return size;
}

void HnswIndex::serialize(char* out) const {
    size_t cnt = cur_element_count_;
    put_raw(out, offset_level0_);
    put_raw(out, max_elements_);
    put_raw(out, cnt);
    put_raw(out, size_data_per_element_);
    put_raw(out, label_offset_);
    put_raw(out, offset_data_);
    put_raw(out, maxlevel_);
    put_raw(out, enterpoint_node_);
    put_raw(out, maxM_);
    put_raw(out, maxM0_);
    put_raw(out, M_);
    put_raw(out, mult_);
    put_raw(out, ef_construction_);
    memcpy(out, level0_, cnt * size_data_per_element_);
    out += cnt * size_data_per_element_;
    for (size_t i = 0; i < cnt; i++) {
        uint32_t sz = element_levels_[i] > 0 ? (uint32_t)(size_links_per_element_ * element_levels_[i]) : 0;
        put_raw(out, sz);
        if (sz) {
            memcpy(out, link_lists_[i], sz);
            out += sz;
        }
    }
}

void HnswIndex::save(const std::string& path) const {
    std::vector<char> buf(serialized_size());
    serialize(buf.data());
    std::ofstream out(path, std::ios::binary);
    if (!out.is_open()) throw std::runtime_error("Cannot open file");
    out.write(buf.data(), (std::streamsize)buf.size());
}

}  // namespace hnswgpu
