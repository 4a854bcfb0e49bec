// CPU oracle: a restatement of the reference hnswlib query path, used ONLY as a checker by the
// tests, by the smoke check and as the timed CPU baseline leg of bench.py.  It is test
// infrastructure and is never linked into, loaded by or called from the product library.
//
// Functions and the reference code each one follows:
//   or_distance            space_l2.h:25-54,97-205,208-238 / space_ip.h:146-204,255-398
//                          (kernel chosen by dim as on an AVX-512 host; GCC contracts the
//                          intrinsic mul+add pairs into FMA, which is restated with fmaf)
//   or_hnsw_load           hnswalg.h:715-821 (on-disk format; header field order 685-712)
//   or_hnsw_search         hnswalg.h:1269-1323 (searchKnn) + 308-440 (searchBaseLayerST<true/false>)
//                          + bindings.cpp:612-697 (row output, closest first)
//   or_hnsw_search_hopcap  the same with a stop after a given number of level-0 hops
//                          (hnswalg.h:343-355 + hnswlib.h:134-150, the stop-condition slot)
//   or_hnsw_search_epsilon hnswalg.h:1326-1377 (searchStopConditionClosest) + stop_condition.h:207-275
//   or_hnsw_search_multivector  the same + stop_condition.h:82-200 (MultiVectorSearchStopCondition)
//   or_bf_search           bruteforce.h:106-125 + bindings.cpp:850-929
//   or_mlp_*               the learned-stop network's forward/backward/Adam in FP64/FP32 scalar
//                          form (no reference counterpart: checks the GPU training kernels)
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_SPACE_L2 0
#define OR_SPACE_IP 1

// Tie modes of the level-0 search:
//   OR_MODE_HEAP  - std::priority_queue ordered by distance only (CompareByFirst,
//                   hnswalg.h:165-170): ties resolve exactly as the reference's heaps do.
//   OR_MODE_TOTAL - the same algorithm with every heap keyed by (distance, internal id):
//                   the deterministic total order the GPU search implements.
#define OR_MODE_HEAP 0
#define OR_MODE_TOTAL 1

const char* or_last_error(void);

float or_distance(int space, size_t dim, const float* a, const float* b);
void or_distance_batch(int space, size_t dim, const float* a, const float* b, size_t n, float* out);
double or_distance_fp64(int space, size_t dim, const float* a, const float* b);

void* or_hnsw_load(const char* path);
void* or_hnsw_load_bytes(const void* buf, size_t len);
void or_hnsw_free(void* h);
size_t or_hnsw_count(void* h);
int or_hnsw_maxlevel(void* h);
uint32_t or_hnsw_entry(void* h);

// Per-query counters written when `counters` is non-null (4 int64 per query):
//   [0] level-0 expansions H0, [1] level-0 distance evaluations E0 (entry included),
//   [2] upper-layer hops Hu, [3] upper-layer distance evaluations Eu (entry included).
// `tie_flags` (optional, one byte per query) is set to 1 when two distances evaluated during
// that query's level-0 search compare equal, i.e. when a tie order could have been consulted.
// ef_per_query may be null (then `ef` applies to all).  k_per_query may be null (then `k`).
// Rows with fewer than k results are padded with label UINT64_MAX and +inf; the return
// value is the number of such rows, or -1 on error.
int or_hnsw_search(void* h, int space, const float* q, size_t nq, size_t k, size_t ef,
                   const int32_t* ef_per_query, const int32_t* k_per_query, size_t k_stride, int mode,
                   int nthreads, uint64_t* labels, float* dists, int64_t* counters, uint8_t* tie_flags);
// The same with a filter: allowed[internal id] != 0 admits the element to the result set (null =
// every element).  With a filter, or with delete marks in the index, the non-bare-bone branch of
// searchBaseLayerST runs (hnswalg.h:323-437; searchKnn's choice, hnswalg.h:1304-1312).
int or_hnsw_search_allowed(void* h, int space, const float* q, size_t nq, size_t k, size_t ef,
                           const int32_t* ef_per_query, const int32_t* k_per_query, size_t k_stride, int mode,
                           int nthreads, const uint8_t* allowed, uint64_t* labels, float* dists, int64_t* counters,
                           uint8_t* tie_flags);
// The bare-bone search with a per-query hop cap: query r also stops once it has expanded
// hop_caps[r] level-0 nodes (hop_caps[r] <= 0: no cap).  This is searchBaseLayerST driven by a
// stop condition that ends the search after a given number of hops (hnswalg.h:343-355 with
// hnswlib.h:134-150), the shape of a learned stop that fired after that many expansions.
int or_hnsw_search_hopcap(void* h, int space, const float* q, size_t nq, size_t k, size_t ef,
                          const int32_t* hop_caps, size_t k_stride, int mode, int nthreads, uint64_t* labels,
                          float* dists, int64_t* counters, uint8_t* tie_flags);
// searchStopConditionClosest with EpsilonSearchStopCondition (hnswalg.h:1326-1377,
// stop_condition.h:207-275): per query, counts[q] results (distance <= eps, at most max_c),
// closest first, in rows of out_stride (the first out_stride of them).  Mode HEAP keeps the
// reference's result order; mode TOTAL orders by (distance, label).  Returns 0, or -1 on error.
int or_hnsw_search_epsilon(void* h, int space, const float* q, size_t nq, float eps, size_t min_c, size_t max_c,
                           int mode, int nthreads, const uint8_t* allowed, size_t out_stride, uint64_t* labels,
                           float* dists, uint32_t* counts, int64_t* counters, uint8_t* tie_flags);
// searchStopConditionClosest with MultiVectorSearchStopCondition (hnswalg.h:1326-1377,
// stop_condition.h:82-200); docs[internal id] = the element's document.  Output as for
// or_hnsw_search_epsilon.
int or_hnsw_search_multivector(void* h, int space, const float* q, size_t nq, const uint32_t* docs,
                               size_t num_docs, size_t ef_collection, int mode, int nthreads, const uint8_t* allowed,
                               size_t out_stride, uint64_t* labels, float* dists, uint32_t* counts,
                               int64_t* counters, uint8_t* tie_flags);
// Internal id of each element's label and its delete mark, for building `allowed` masks.
void or_hnsw_labels(void* h, uint64_t* labels_by_id, uint8_t* deleted_by_id);

// Exact brute force: the k lexicographically smallest (distance, label) pairs per query,
// closest first.  labels_in may be null (labels = row index).
int or_bf_search(int space, size_t dim, const float* data, const uint64_t* labels_in, size_t n,
                 const float* q, size_t nq, size_t k, int nthreads, uint64_t* labels, float* dists);

// FP64 distances for listed (query, row) pairs: out[i] = d(q[qi[i]], data[ri[i]]).
void or_fp64_pairs(int space, size_t dim, const float* q, const float* data, const int64_t* qi,
                   const int64_t* ri, size_t n, double* out);

// Learned-stop MLP (in -> hidden -> hidden -> 1, GELU(tanh form)), row-major weights:
// W1[h][in], b1[h], W2[h][h], b2[h], w3[h], b3.  Parameters are FP32; math in FP64.
void or_mlp_forward(const float* params, int in_dim, int hidden, const float* x, size_t n, float* logits);
// Mean binary-cross-entropy-with-logits loss and its gradient w.r.t. every parameter.
double or_mlp_loss_grad(const float* params, int in_dim, int hidden, const float* x, const float* y,
                        size_t n, double* grad);
int or_mlp_param_count(int in_dim, int hidden);

#ifdef __cplusplus
}
#endif
