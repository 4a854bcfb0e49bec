// Shared device-side definitions for the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hg {

constexpr uint32_t SENT = 0xFFFFFFFFu;    // empty link slot / empty hash slot / no result
constexpr uint32_t XFLAG = 0x80000000u;   // "expanded" mark carried in a beam id
constexpr uint32_t DFLAG = 0x40000000u;   // "not in the result set" (deleted or filtered out)
constexpr uint32_t ID_MASK = 0x3FFFFFFFu; // ids below 2^30

// GPU layout of an HNSW graph (exported once from the CPU index):
//   vec      [n][stride] f32, stride = dim rounded up to 4 (16-byte rows), zero padded
//   links0   [n][maxM0] u32 level-0 neighbours in list order, SENT padded
//   upper_off[n] u32: row of level 1 in `upper` (levels l>1 follow at +l-1), SENT if none
//   upper    [rows][maxM] u32 upper-level neighbours, SENT padded
//   labels   [n] u64 external labels
//   vec8     optional [n][dim] u8: the same vectors when every stored coordinate is an integer in
//            0..255 (its FP32 bits those of the byte's value), dim % 16 == 0.  Row r holds, for
//            lane g = 0..3 of a distance group, the words t = 0..dim/16-1 of coordinates
//            16t+4g..16t+4g+3: byte (g * dim/16 + t) * 4 + j = coordinate 16t + 4g + j, so each
//            lane reads one contiguous run.  The query kernel reads it instead of vec when set.
struct DevGraph {
    const float* vec;
    const uint8_t* vec8;
    const uint32_t* links0;
    const uint32_t* upper_off;
    const uint32_t* upper;
    const uint64_t* labels;
    uint32_t n, dim, stride, maxM, maxM0;
    int32_t maxlevel;
    uint32_t entry;
    int32_t space;
};

// Learned early-termination network held in device memory (FP32 master copy plus BF16
// weights used by the in-kernel inference).
struct DevStopModel {
    const float* params;        // W1[h][in] b1[h] W2[h][h] b2[h] w3[h] b3
    const float* mean;          // [in]
    const float* inv_std;       // [in]
    int32_t in_dim, hidden;
    int32_t period;             // evaluate every `period` level-0 expansions
    int32_t min_hops;           // never stop before this many expansions
    float logit_threshold;      // stop when logit >= threshold (= logit(confidence))
    float ext_logit_threshold;  // at the search's own end: extend when logit < this
};

struct SearchParams {
    DevGraph g;
    const float* queries;       // [nq][dim]
    const uint32_t* order;      // optional permutation: slot -> query index
    uint32_t nq;
    const int32_t* ef_pq;       // optional per-query ef
    const int32_t* k_pq;        // optional per-query k
    uint32_t ef, k, k_stride;
    uint64_t* out_labels;       // [nq][k_stride]
    float* out_dists;           // [nq][k_stride]
    int32_t* counters;          // optional [nq][4]
    uint32_t* work;             // persistent work counter (zeroed before each launch)
    uint32_t cap;               // beam capacity per wave (multiple of 64, >= every max(ef,k))
    uint32_t hash_bits;         // visited hash has 1 << hash_bits slots per wave
    // learned stop (optional)
    DevStopModel stop;
    int32_t use_stop;
    // trace collection (optional)
    float* trace_feat;          // [nq][max_rec][in_dim]
    uint32_t* trace_topk;       // [nq][max_rec][k]
    int32_t* trace_nrec;        // [nq]
    uint32_t* trace_final;      // [nq][k] internal ids of the final top-k
    int32_t trace_period, trace_max_rec;
    // non-bare-bone search (deletions / label filter): bitmaps over internal ids, bit set =
    // deleted (del_bits) or allowed (allow_bits); either may be null.  status bit 0 is set when
    // a query needed more beam entries than cap (its result is then not exact).
    const uint32_t* del_bits;
    const uint32_t* allow_bits;
    uint32_t* status;
    // FILT, first pass: a query that outgrew the beam appends its index to ovf_list (counter
    // ovf_count) instead of setting status; the second pass takes its queries from that list
    // (order = ovf_list, nq_dev = ovf_count: the count is read on the device).
    uint32_t* ovf_list;
    uint32_t* ovf_count;
    const uint32_t* nq_dev;
    // FILT, range search (EpsilonSearchStopCondition): ef = the result bound max_c; the search
    // also stops at a candidate farther than eps once eps_min results are held; results are
    // those within eps, counts[q] of them, the first k_stride written.
    int eps_mode;
    float eps;
    int eps_min;
    uint32_t* counts;
    // graph construction: the upper-level descent ignores ids >= id_limit (elements not yet in
    // the level-0 graph; 0 = no limit); queries are rows of q_stride floats (0 = dim)
    uint32_t id_limit;
    uint32_t q_stride;
    // optional per-query cap on level-0 expansions: query qi stops once it has expanded
    // hop_cap[qi] nodes (a value <= 0 means no cap)
    const int32_t* hop_cap;
    // learned stop, first pass of an extended search: a query whose search ends by its own rule
    // while the network is not confident its top-k has converged is appended to ext_list
    // (counter ext_count) with no output; a later launch at the larger ef answers it
    uint32_t* ext_list;
    uint32_t* ext_count;
    // trace collection: also record the features where the search ends by its own rule
    int32_t trace_end;
};

// Multi-vector search (MultiVectorSearchStopCondition): one wave per query; every element
// belongs to a document docs[id]; the result set is bounded by its number of distinct documents.
struct MvParams {
    DevGraph g;
    const float* queries;  // [nq][dim]
    uint32_t nq;
    uint32_t* work;        // query counter (zeroed by the launch)
    const uint32_t* docs;  // [n]
    const uint32_t* del_bits;    // optional
    const uint32_t* allow_bits;  // optional
    uint32_t num_docs, ef_coll;  // ef_coll >= num_docs
    uint32_t cap, hash_bits, dcap;
    uint64_t* out_labels;  // [nq][k_stride]
    float* out_dists;
    uint32_t* counts;      // [nq]
    uint32_t k_stride;
    int32_t* counters;     // optional [nq][4]
    uint32_t* status;      // bit 0: a query outgrew the beam
};

constexpr int kStopFeatures = 12;

}  // namespace hg
