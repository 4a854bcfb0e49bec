/* hnswgpu: MI355X (gfx950) query engine for hnswlib indexes -- C ABI.
 *
 * Plain pointers and sizes only.  Every call returns HG_OK (0) or an error code; the message
 * of the last error on the calling thread is returned by hg_last_error().  Messages reuse the
 * reference's texts where the condition is the same one.
 *
 * Which reference interface each entry point replaces (file:line in hnswlib v0.9.0):
 *   hg_hnsw_create        HierarchicalNSW(space, max_elements, M, ef_construction, random_seed)
 *                         hnswalg.h:89-144; Index.init_index bindings.cpp:192-207
 *   hg_hnsw_add_items     Index.add_items -> ParallelFor -> addPoint  bindings.cpp:251-304,
 *                         hnswalg.h:1152-1266 (CPU: the index stays "built by addPoint")
 *   hg_hnsw_build_gpu     Index.add_items -> addPoint hnswalg.h:1152-1266 with the level-0
 *                         links built on the GPU in batches (a graph of the same kind, not
 *                         the same graph; the levels above 0 are the CPU's)
 *   hg_hnsw_save / _load  saveIndex / loadIndex  hnswalg.h:685-821
 *   hg_hnsw_serialize     the same byte stream as saveIndex, into memory
 *   hg_gpu_index_create   (new) export of data_level0_memory_ / linkLists_ into the GPU layout
 *   hg_gpu_index_save / _load  (new) versioned cache of that layout
 *   hg_gpu_search         Index.knn_query bindings.cpp:612-697 over
 *                         HierarchicalNSW::searchKnn hnswalg.h:1269-1323 (bare-bone branch;
 *                         the non-bare-bone branch when the index holds deleted elements)
 *   hg_gpu_search_filtered  the same with a filter: knn_query(filter=) bindings.cpp:612-697 ->
 *                         searchKnn(query, k, isIdAllowed), non-bare-bone searchBaseLayerST
 *                         hnswalg.h:308-440
 *   hg_gpu_search_device  the same with device-resident queries/results, per-query ef and k
 *                         (extends the single ef_ of hnswalg.h:173-175)
 *   hg_gpu_search_epsilon searchStopConditionClosest(q, EpsilonSearchStopCondition, filter)
 *                         hnswalg.h:1326-1377, stop_condition.h:207-275 (range search)
 *   hg_gpu_search_multivector  searchStopConditionClosest(q, MultiVectorSearchStopCondition)
 *                         hnswalg.h:1326-1377, stop_condition.h:82-200 (documents of vectors)
 *   hg_bf_create / hg_bf_add   BruteforceSearch(space, maxElements) / addPoint
 *                         bruteforce.h:46-83; BFIndex.init_index/add_items bindings.cpp:785-819
 *   hg_bf_search          BFIndex.knn_query bindings.cpp:850-929 over
 *                         BruteforceSearch::searchKnn bruteforce.h:106-125
 *   hg_bf_search_device   the same, device-resident, with reduced-precision distance GEMM
 *                         (BF16 / FP8) and an exact FP32 re-rank of the top-c candidates
 *   hg_bf_search_precision  BFIndex.knn_query bindings.cpp:850-929 with the certified
 *                         reduced-precision search (same lists as bruteforce.h:106-125)
 *   hg_bf_search_filtered BFIndex.knn_query(filter=) -> searchKnn(q, k, isIdAllowed)
 *                         bruteforce.h:106-125 (exact FP32)
 *   hg_bf_delete          BFIndex.delete_vector bindings.cpp:829-831 -> removePoint
 *                         bruteforce.h:85-104
 *   hg_bf_save / _load    BFIndex.save_index / load_index -> saveIndex / loadIndex
 *                         bruteforce.h:128-161
 */
#ifndef HNSWGPU_H
#define HNSWGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HG_OK 0
#define HG_ERR_INVALID 1  /* bad argument (reference: std::runtime_error on the same checks) */
#define HG_ERR_RUNTIME 2  /* index/algorithm error (capacity, corrupted file, cand error ...) */
#define HG_ERR_HIP 3      /* HIP runtime failure (device OOM, launch failure ...) */
#define HG_ERR_SHORT 4    /* some query found fewer than k results */

#define HG_SPACE_L2 0
#define HG_SPACE_IP 1

#define HG_PREC_FP32 0 /* exact: reference distance order, no candidates dropped */
#define HG_PREC_BF16 1 /* BF16 MFMA distance GEMM, FP32 accumulate, top-c FP32 re-rank */
#define HG_PREC_FP8 2  /* FP8 e4m3 (per-vector scale) MFMA GEMM, FP32 accumulate, re-rank */

/* Vector rows the GPU graph search reads (hg_gpu_set_vector_code / hg_gpu_vector_code). */
#define HG_VEC_F32 0  /* the FP32 rows as stored by the index */
#define HG_VEC_U8 1   /* an 8-bit copy: every coordinate is an integer in 0..255 (lossless) */
#define HG_VEC_AUTO 1 /* the 8-bit copy where the export found it lossless, else FP32 */

const char* hg_last_error(void);

/* ---- CPU-side HNSW index (construction prerequisite, persistence) ---- */
int hg_hnsw_create(int space, size_t dim, size_t max_elements, size_t M, size_t ef_construction,
                   size_t random_seed, void** out_index);
int hg_hnsw_load(int space, size_t dim, const char* path, size_t max_elements, void** out_index);
/* labels may be NULL (labels continue from the index's running counter). */
int hg_hnsw_add_items(void* index, const float* data, size_t rows, const uint64_t* labels, int num_threads);
int hg_hnsw_save(void* index, const char* path);
/* Writes at most `capacity` bytes; *size receives the full serialized size. */
int hg_hnsw_serialize(void* index, void* buffer, size_t capacity, size_t* size);
/* info[0..7] = count, max_elements, dim, M, ef_construction, maxlevel, entry, num_deleted */
int hg_hnsw_info(void* index, int64_t* info);
int hg_hnsw_set_ef(void* index, size_t ef);
/* HierarchicalNSW::markDelete / unmarkDelete (hnswalg.h:852-927): set / clear the delete mark of
 * a label; deleted elements stay in the graph, are traversed, and are never returned.  Errors
 * (HG_ERR_RUNTIME): "Label not found", "The requested to delete element is already deleted",
 * "The requested to undelete element is not deleted". */
int hg_hnsw_mark_deleted(void* index, uint64_t label);
int hg_hnsw_unmark_deleted(void* index, uint64_t label);
int hg_hnsw_get_ef(void* index, size_t* ef);
int hg_hnsw_free(void* index);
/* add_items with the graph built on the GPU (device = HIP device).  Level 0 is built in batches
 * of up to 16,384 rows that grow with the graph: a beam search with ef_construction over the
 * elements before the batch, the earlier rows of the same batch added to the candidates by brute
 * force, the neighbour heuristic, and mutual linking with re-selection of full lists
 * (hnswalg.h:225-305, 443-483, 506-630).  The levels above 0 do not depend on level 0: they are
 * inserted on the CPU (num_threads as hg_hnsw_add_items), or, for a build of an empty index of
 * at least hg_hnsw_build_gpu_upper_min_rows rows, built the same way on the GPU as the level 0
 * of the HNSW that the elements of level >= 1 form one level down (lists of maxM = M links, the
 * reference's rule for those levels).  The index must hold no deleted elements;
 * ef_construction <= 1024.  If the call fails part way, discard the index. */
int hg_hnsw_build_gpu(void* index, const float* data, size_t rows, const uint64_t* labels, int num_threads,
                      int device);
/* Row count from which hg_hnsw_build_gpu builds the levels above 0 on the GPU too (default
 * 250,000).  No counterpart in the reference: a tuning knob of the GPU build. */
int hg_hnsw_build_gpu_upper_min_rows(size_t rows);
/* Its first step alone: insert rows into the levels above 0 (level-0 lists left empty). */
int hg_hnsw_add_items_upper(void* index, const float* data, size_t rows, const uint64_t* labels, int num_threads);
/* Index.normalize_vector (bindings.cpp:241-248) applied row by row (the cosine space). */
int hg_normalize_rows(const float* in, float* out, size_t rows, size_t dim);

/* ---- GPU copy of an HNSW index ---- */
int hg_gpu_index_create(void* index, int device, void** out_gpu_index);
/* Versioned cache of the GPU layout (no counterpart in the reference: the exported arrays, delete
 * marks included, with a format version and a checksum), so that a process can put an index on
 * the device without the CPU index: hg_gpu_index_save writes it, hg_gpu_index_load checks the
 * version, checksum and ids and uploads it.  A file of another version is refused. */
int hg_gpu_index_save(void* gpu_index, const char* path);
int hg_gpu_index_load(const char* path, int device, void** out_gpu_index);
int hg_gpu_index_free(void* gpu_index);
/* Bytes resident on the device for this index. */
int hg_gpu_index_bytes(void* gpu_index, size_t* bytes);
/* No counterpart in the reference (a storage choice of the GPU layout; the reference keeps FP32
 * rows in data_level0_memory_, hnswalg.h:121-127).  At export, when every stored coordinate is an
 * integer in 0..255 whose FP32 bits are the stored ones (descriptor data such as SIFT), dim is
 * 16, 32, 64, 96 or 128 and the rows are exact, the export also keeps an 8-bit copy of the
 * vectors, and the graph search reads it (a quarter of the bytes per distance evaluation); the
 * kernel converts each byte back to that same FP32 value, so every distance is bit-identical to
 * the FP32 rows'.  HG_VEC_AUTO (default) uses the copy where it exists; HG_VEC_F32 makes the
 * search read the FP32 rows.  hg_gpu_vector_code reports which rows the search reads
 * (HG_VEC_U8 or HG_VEC_F32).  Changing the choice drops the autotuned launch plans; searches
 * already launched and graphs captured before the change keep the rows they were launched with
 * (both copies stay resident, and both give the same results). */
int hg_gpu_set_vector_code(void* gpu_index, int code);
int hg_gpu_vector_code(void* gpu_index, int* code);

/* knn_query on host buffers: queries [nq][dim] f32; labels/dists [nq][k], closest first.
 * counters (optional) receives 4 int32 per query: level-0 expansions, level-0 distance
 * evaluations (entry included), upper-layer hops, upper-layer evaluations (entry included).
 * Returns HG_ERR_SHORT (after writing every row, short rows padded with UINT64_MAX / +inf)
 * when a query finds fewer than k results. */
int hg_gpu_search(void* gpu_index, const float* queries, size_t nq, size_t k, size_t ef, uint64_t* labels,
                  float* dists, int32_t* counters);

/* Device-resident batched search.  ef_per_query / k_per_query (device int32 [nq]) may be NULL,
 * then ef / k apply to all queries; when given, ef and k must be at least every per-query value
 * (they size the launch: a larger per-query k is clamped to k, a larger max(ef, k) to the beam
 * capacity).  Rows of labels/dists have stride k_stride (>= every k).
 * stop_model (may be NULL) enables learned early termination (hg_stop_model_*).
 * stream is a hipStream_t (NULL = default stream).  Asynchronous: no host synchronisation. */
int hg_gpu_search_device(void* gpu_index, const float* d_queries, size_t nq, const int32_t* d_ef_per_query,
                         const int32_t* d_k_per_query, size_t ef, size_t k, size_t k_stride, uint64_t* d_labels,
                         float* d_dists, int32_t* d_counters, void* stop_model, void* stream);

/* Filtered search (knn_query(..., filter=) bindings.cpp:612-697; BaseFilterFunctor
 * hnswlib.h:16-20 evaluated by the non-bare-bone searchBaseLayerST hnswalg.h:308-440).
 * A filter is built once per GPU index from the labels it allows (labels absent from the index
 * are ignored) and may be reused by any number of searches.  Deleted elements (markDelete) are
 * never returned, filter or not.  The beam holds every admitted element until ef allowed ones
 * are found, so a very selective filter needs a large beam: the host search retries at the
 * largest beam (8192 entries) and fails with HG_ERR_RUNTIME if a query still needs more. */
int hg_gpu_filter_create(void* gpu_index, const uint64_t* allowed_labels, size_t n_allowed, void** out_filter);
int hg_gpu_filter_free(void* filter);
/* Elements of the index the filter allows. */
int hg_gpu_filter_count(void* filter, size_t* allowed);
/* hg_gpu_search with a filter (NULL = no filter). */
int hg_gpu_search_filtered(void* gpu_index, const float* queries, size_t nq, size_t k, size_t ef, void* filter,
                           uint64_t* labels, float* dists, int32_t* counters);
/* Device-resident filtered search (as hg_gpu_search_device, no stop model).  Two launches on
 * the stream: the beam planned from the allowed fraction, then, for the queries that outgrew
 * it (listed on the device), the largest beam.  d_status (device uint32, may be NULL = the
 * index's own word) is cleared by the call and gets bit 0 set when a query outgrew even the
 * largest beam; its results are then not those of the reference. */
int hg_gpu_search_device_filtered(void* gpu_index, const float* d_queries, size_t nq, const int32_t* d_ef_per_query,
                                  const int32_t* d_k_per_query, size_t ef, size_t k, size_t k_stride,
                                  uint64_t* d_labels, float* d_dists, int32_t* d_counters, void* filter,
                                  uint32_t* d_status, void* stream);
/* The bare-bone device search stopped after d_hop_cap[q] level-0 expansions (device int32 [nq];
 * <= 0 = no cap): searchBaseLayerST driven by a stop condition that ends the search after a given
 * number of hops (hnswalg.h:343-355 with hnswlib.h:134-150).  It replays a learned search's
 * stopping points without the network (to time the two hop for hop) and is the GPU side of the
 * oracle's hop-capped mode.  Asynchronous on stream. */
int hg_gpu_search_device_hopcap(void* gpu_index, const float* d_queries, size_t nq, const int32_t* d_hop_cap,
                                size_t ef, size_t k, uint64_t* d_labels, float* d_dists, int32_t* d_counters,
                                void* stream);
/* hg_gpu_search_device_filtered (filter may be NULL; the bare-bone search when the index holds no
 * deleted elements and no filter is given) that allocates nothing: no hipMalloc, hipFree or other
 * allocating call.  No counterpart in the reference: it is for callers that cannot allocate --
 * code under hipGraph capture, tight serving loops, harnesses that forbid allocation.  Its device
 * words (work counters, status, the non-bare-bone search's overflow lists) live in one fixed
 * device array per GPU (2 MiB + 256 B, held until the process ends), so with deleted elements or
 * a filter a call takes at most 262,144 queries (HG_ERR_INVALID beyond).  Every call on one GPU
 * shares that array: calls that overlap on one GPU (other streams, other threads) race; the
 * array keeps the last call's contents and each call clears what it reads.  d_status as for
 * hg_gpu_search_device_filtered (NULL = the array's own word, not readable by the caller). */
int hg_gpu_search_device_noalloc(void* gpu_index, const float* d_queries, size_t nq, const int32_t* d_ef_per_query,
                                 const int32_t* d_k_per_query, size_t ef, size_t k, size_t k_stride,
                                 uint64_t* d_labels, float* d_dists, int32_t* d_counters, void* filter,
                                 uint32_t* d_status, void* stream);
/* The index's own status word (see above) and the number of queries that needed the second,
 * largest-beam pass, for the last search launched on this index; waits for that launch to end.
 * Either pointer may be NULL. */
int hg_gpu_search_status(void* gpu_index, uint32_t* status, uint32_t* second_pass_queries);
/* Range search: searchStopConditionClosest with EpsilonSearchStopCondition(epsilon,
 * min_candidates, max_candidates) (hnswalg.h:1326-1377, stop_condition.h:207-275), optionally
 * filtered (filter may be NULL; deleted elements are never returned).  epsilon bounds the
 * distance as the space reports it (squared L2, or 1 - inner product).  counts[q] = number of
 * results of query q (at most max_candidates); row q of labels/dists (stride out_stride <= 8192)
 * holds the closest min(counts[q], out_stride) of them, ordered by (distance, label), padded
 * with UINT64_MAX / +inf.  The reference returns internal ids in place of labels here; the two
 * agree when labels are the insertion order.  Fails (HG_ERR_RUNTIME) when a query's beam needs
 * more than 8192 entries. */
int hg_gpu_search_epsilon(void* gpu_index, const float* queries, size_t nq, float epsilon, size_t min_candidates,
                          size_t max_candidates, void* filter, size_t out_stride, uint64_t* labels, float* dists,
                          uint32_t* counts);
/* The same on device buffers, asynchronous on stream; d_counters and d_status as for
 * hg_gpu_search_device_filtered (each may be NULL). */
int hg_gpu_search_epsilon_device(void* gpu_index, const float* d_queries, size_t nq, float epsilon,
                                 size_t min_candidates, size_t max_candidates, void* filter, size_t out_stride,
                                 uint64_t* d_labels, float* d_dists, uint32_t* d_counts, int32_t* d_counters,
                                 uint32_t* d_status, void* stream);
/* Multi-vector search: searchStopConditionClosest with MultiVectorSearchStopCondition(space,
 * num_docs, ef_collection) (hnswalg.h:1326-1377, stop_condition.h:82-200).  The reference keeps
 * each element's document id in its vector record (MultiVectorL2Space /
 * MultiVectorInnerProductSpace); here a documents object maps every element of the GPU index
 * to its document, built from (label, document) pairs (every element needs one).  The result
 * set holds at most max(ef_collection, num_docs) documents during the search and num_docs at
 * the end; counts[q] results per query, rows as hg_gpu_search_epsilon.  The distances read the
 * vector only, so an index built from the vectors alone is the reference's multi-vector graph.
 * ef_collection <= 1024; the host call retries a batch at an 8192-entry beam when a query
 * outgrows 2048 entries and then fails (HG_ERR_RUNTIME). */
int hg_gpu_docs_create(void* gpu_index, const uint64_t* labels, const uint32_t* docs, size_t count, void** out_docs);
int hg_gpu_docs_free(void* docs);
int hg_gpu_search_multivector(void* gpu_index, const float* queries, size_t nq, void* docs, size_t num_docs,
                              size_t ef_collection, void* filter, size_t out_stride, uint64_t* labels, float* dists,
                              uint32_t* counts);
/* The same on device buffers at one beam capacity (cap, 0 = 2048); asynchronous on stream;
 * d_status (required) bit 0 reports a query that outgrew the beam. */
int hg_gpu_search_multivector_device(void* gpu_index, const float* d_queries, size_t nq, void* docs,
                                     size_t num_docs, size_t ef_collection, void* filter, size_t out_stride,
                                     uint64_t* d_labels, float* d_dists, uint32_t* d_counts, int32_t* d_counters,
                                     uint32_t* d_status, uint32_t cap, void* stream);
/* Re-reads the delete marks (markDelete / unmarkDelete) of the CPU index this GPU index was
 * exported from, without a new export.  Fails if elements were added since the export. */
int hg_gpu_refresh_deleted(void* gpu_index, void* index);
/* Labels by internal id (getExternalLabel hnswalg.h:185-189 for every element); labels may be
 * NULL to query the count. */
int hg_gpu_index_labels(void* gpu_index, uint64_t* labels, size_t capacity, size_t* count);

/* Launch configuration chosen by the autotuner for (max ef, max k) on this index:
 * out[0] = waves per workgroup, out[1] = visited-hash slots per wave, out[2] = beam capacity,
 * out[3] = LDS bytes per workgroup, out[4] = workgroups launched. */
int hg_gpu_search_plan(void* gpu_index, size_t max_ef, size_t max_k, size_t nq, int32_t* out);
/* Re-tune: time candidate (visited-hash slots 2^7..2^13, workgroups per CU) pairs on the given
 * device-resident queries and keep the fastest for this (max ef, max k) bucket. */
int hg_gpu_autotune(void* gpu_index, const float* d_queries, size_t nq, size_t ef, size_t k, void* stream);
/* Stores a launch plan (visited-hash slots = 1 << hash_bits, persistent workgroups per CU) for
 * the beam capacity of (ef, k), in place of an autotuned one.  No counterpart in the reference:
 * a tuning knob of the GPU search for callers that already know their plan.  hash_bits in
 * [7, 16], workgroups_per_cu in [1, 32]. */
int hg_gpu_set_plan(void* gpu_index, size_t ef, size_t k, uint32_t hash_bits, int workgroups_per_cu);

/* ---- Brute force ---- */
int hg_bf_create(int space, size_t dim, size_t max_elements, int device, void** out_bf);
/* labels may be NULL (row index continues from the current count). */
int hg_bf_add(void* bf, const float* data, size_t rows, const uint64_t* labels);
int hg_bf_count(void* bf, size_t* count);
int hg_bf_free(void* bf);
/* Exact FP32 BFIndex.knn_query on host buffers (closest first, ties by label). */
int hg_bf_search(void* bf, const float* queries, size_t nq, size_t k, uint64_t* labels, float* dists);
/* Device-resident search.  precision = HG_PREC_*; rerank_c (<= 4096) = candidates kept per
 * query by the reduced-precision pass and re-ranked in exact FP32 (ignored for FP32).
 * Reduced precision is certified per query: with B a bound on |approximate - exact FP32
 * distance| over all rows (from the query's and the rows' norms and quantisation errors and
 * the FP32 rounding of both forms), a query whose rows left out of the re-rank all have
 * approximate distance >= A, where A - B exceeds the k-th exact distance found, has the exact
 * search's list; every other query (and one whose candidate list overflowed) is answered by the
 * exact path.  The lists therefore always equal BFIndex.knn_query's (bruteforce.h:106-125).
 * Adjacent near-tie pairs of the re-ranked order (FP32 distances within 1e-6 relative, up to
 * the pair at the k-th boundary) are recomputed in FP64 (hg_bf_lowp_report).  stats (optional,
 * host, 4 doubles): GEMM ms, select ms, re-rank ms, queries answered by the exact path.
 * Synchronises the stream (the exact-path queries are chosen on the host). */
int hg_bf_search_device(void* bf, const float* d_queries, size_t nq, size_t k, int precision, size_t rerank_c,
                        uint64_t* d_labels, float* d_dists, double* stats, void* stream);
/* BFIndex.knn_query (bindings.cpp:850-929) on host buffers at a chosen precision: HG_PREC_FP32
 * is hg_bf_search; HG_PREC_BF16 / HG_PREC_FP8 run the certified reduced-precision search above
 * (dim % 64 == 0), whose lists equal hg_bf_search's. */
int hg_bf_search_precision(void* bf, const float* queries, size_t nq, size_t k, int precision, size_t rerank_c,
                           uint64_t* labels, float* dists);
/* What the last reduced-precision search on this index did besides the re-rank: out[0] queries
 * answered by the exact path because the candidate list overflowed, came up short or had too
 * many ties at the cut, out[1] queries answered by the exact path because the bound could not
 * certify them, out[2] near-tie pairs checked in FP64, out[3] those whose FP64 order differs from
 * the FP32 order (the reference's order is the FP32 one: the lists keep it), out[4] / out[5] the
 * same two counts for the pair at the k-th boundary, out[6] the largest per-query bound, out[7]
 * 1 when the bound was applied, out[8] rows whose exact FP32 distance the re-rank computed (with
 * the bound, only the rows whose approximate distance is within twice the bound of the (k+1)-th
 * smallest among the candidates).  No counterpart in the reference. */
int hg_bf_lowp_report(void* bf, double* out);
/* FP64 report on final lists (any precision): every listed (query, label) distance is recomputed
 * in FP64; pairs of neighbouring results whose FP32 distances differ by at most rel_tol
 * (relative) are near ties.  out[0] = near-tie pairs, out[1] = those whose FP64 order disagrees
 * with the list's (distance, label) order, out[2] = max |d32 - d64| over all returned results,
 * out[3] = max relative error.  Labels are mapped to rows through the index's label table; a
 * label the index does not hold is an error. */
int hg_bf_fp64_check(void* bf, const float* d_queries, size_t nq, size_t k, const uint64_t* d_labels,
                     const float* d_dists, double rel_tol, double* out, void* stream);

/* Label of every row, in row order (capacity >= the count). */
int hg_bf_labels(void* bf, uint64_t* labels, size_t capacity);
/* removePoint: the last row moves into the removed row; an unknown label is ignored. */
int hg_bf_delete(void* bf, uint64_t label);
/* saveIndex format: size_t max_elements, size_t bytes per element (dim*4 + 8), size_t count,
 * then max_elements records [dim f32 | u64 label] (records past the count are zero). */
int hg_bf_save(void* bf, const char* path);
/* loadIndex into a new GPU index (capacity from the file); the label -> row map is rebuilt, so
 * later adds of a known label overwrite its row and deletions find it. */
int hg_bf_load(int space, size_t dim, const char* path, int device, void** out_bf);
/* Label filter of a brute-force index (bitmap over its rows; labels absent from the index are
 * ignored).  Valid until the next add or delete on the index. */
int hg_bf_filter_create(void* bf, const uint64_t* allowed_labels, size_t n_allowed, void** out_filter);
int hg_bf_filter_free(void* filter);
/* Exact FP32 search among the allowed rows (filter may be NULL): the k smallest
 * (distance, label) pairs, closest first.  HG_ERR_SHORT when fewer than k rows are allowed. */
int hg_bf_search_filtered(void* bf, const float* queries, size_t nq, size_t k, void* filter, uint64_t* labels,
                          float* dists);
/* The same on device buffers, asynchronous on stream. */
int hg_bf_search_device_filtered(void* bf, const float* d_queries, size_t nq, size_t k, void* filter,
                                 uint64_t* d_labels, float* d_dists, void* stream);

/* Tile configuration of the reduced-precision distance GEMM (process-wide): 0 = 128x128
 * workgroup tiles with 2x2 MFMA blocks per wave and double-buffered LDS; 1 = 256x128 tiles
 * with 4x2 blocks per wave and a register-prefetched LDS stage; 2 (default) = 256x128 tiles
 * fed by direct-to-LDS loads through a ring of three k-slices; 3 = 256x256 tiles, eight waves,
 * the same direct-to-LDS ring; 4 = 256x256 tiles, four waves of 4x4 blocks, the same ring. */
#define HG_GEMM_VARIANTS 5
int hg_bf_gemm_variant(int variant);
/* Autotuner of the GEMM tile configuration for this index and precision (BF16 or FP8): runs the
 * reduced-precision search on the device-resident queries with every variant, keeps the one
 * with the shortest filter GEMM for later hg_bf_search_device calls on this index (it overrides
 * hg_bf_gemm_variant there), writes the GEMM time (ms) of every variant to
 * times_ms[HG_GEMM_VARIANTS] and the choice to
 * *chosen (either may be NULL).  dl/dd receive the last run's results ([nq][k]).  No counterpart
 * in the reference: the tile shape is a GPU concern. */
int hg_bf_autotune(void* bf_index, const float* d_queries, size_t nq, size_t k, int precision, size_t rerank_c,
                   uint64_t* d_labels, float* d_dists, double* times_ms, int* chosen, void* stream);

/* ---- Learned early termination (stop model) ---- */
/* Create a stop model for a given feature count / hidden width (weights zero). */
int hg_stop_model_create(int in_dim, int hidden, int device, void** out_model);
int hg_stop_model_free(void* model);
int hg_stop_model_param_count(void* model, size_t* count);
/* Parameters in FP32 (host), layout W1[h][in], b1[h], W2[h][h], b2[h], w3[h], b3. */
int hg_stop_model_set_params(void* model, const float* params, size_t count);
int hg_stop_model_get_params(void* model, float* params, size_t count);
/* Feature normalisation (mean, 1/std per feature) applied before the network. */
int hg_stop_model_set_norm(void* model, const float* mean, const float* inv_std, size_t count);
/* Inference settings: every `period` level-0 expansions (from min_hops on) the network runs on
 * the 12 traversal features and a query stops when sigmoid(logit) >= threshold, i.e. when the
 * network is confident its top-k has converged; otherwise the search runs to its own end (the
 * plain fixed-ef rule, hnswalg.h:343-355). */
int hg_stop_model_config(void* model, int period, float threshold, int min_hops);
/* Extension: where a query's search ends by its own rule, the network runs once more; when
 * sigmoid(logit) < ext_threshold (not confident the top-k has converged) the query is listed on
 * the device and answered by a second launch at ext_ef -- the plain fixed-ef search at the larger
 * ef, with the same stop checks.  ext_ef = 0 (default) turns it off; it applies when ext_ef >
 * max(ef, k) of the search.  hg_gpu_search_status reports how many queries were extended.  No
 * counterpart in the reference: a stop condition of this project (hnswlib.h:134-150 slot). */
int hg_stop_model_config_extend(void* model, int ext_ef, float ext_threshold);
/* Trace collection: run the plain search and write, per query, up to max_records feature
 * vectors (every `period` level-0 expansions) plus the current top-k internal ids at each
 * record, and the final top-k internal ids; with end_record != 0 one more record where the
 * search ends by its own rule.  d_features [nq][max_records][12] f32, d_topk
 * [nq][max_records][k] u32, d_nrec [nq] int32, d_final [nq][k] u32. */
int hg_gpu_collect_traces(void* gpu_index, const float* d_queries, size_t nq, size_t ef, size_t k, int period,
                          int max_records, float* d_features, uint32_t* d_topk, int32_t* d_nrec, uint32_t* d_final,
                          int end_record, void* stream);
/* External labels of n internal ids (as in trace records and d_final), on the device; an empty
 * slot (0xFFFFFFFF) maps to all ones.  Counterpart of the reference's getExternalLabel
 * (hnswalg.h:185-189) for device arrays. */
int hg_gpu_id_labels(void* gpu_index, const uint32_t* d_ids, size_t n, uint64_t* d_labels, void* stream);
/* Forward pass (logits) on device features [n][in_dim]. */
int hg_stop_model_forward(void* model, const float* d_x, size_t n, float* d_logits, void* stream);
/* One Adam step on a minibatch (mean BCE-with-logits); returns the loss through *loss
 * (host). d_grad_out (optional, device f32 [param_count]) receives the gradient. */
int hg_stop_model_train_step(void* model, const float* d_x, const float* d_y, size_t n, float lr, float beta1,
                             float beta2, float eps, float* loss, float* d_grad_out, void* stream);

/* ---- Device plumbing (so hosts need no other GPU runtime binding) ---- */
int hg_device_count(int* count);
int hg_set_device(int device);
int hg_malloc(void** ptr, size_t bytes);
int hg_free(void* ptr);
int hg_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
/* Waits for all work on the device before copying (results may come from any stream). */
int hg_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int hg_memset(void* dst, int value, size_t bytes, void* stream);
/* Device-to-device copy by a plain streaming kernel (bytes % 16 == 0); also the calibration
 * kernel of known traffic for the FETCH_SIZE / WRITE_SIZE counters. */
int hg_copy_kernel(void* dst, const void* src, size_t bytes, void* stream);
int hg_stream_create(void** stream);
int hg_stream_destroy(void* stream);
int hg_stream_synchronize(void* stream);
int hg_device_synchronize(void);
/* Event timer on a stream: hg_timer_start records, hg_timer_stop records and waits, returning
 * the elapsed milliseconds between the two events. */
int hg_timer_create(void** timer);
int hg_timer_destroy(void* timer);
int hg_timer_start(void* timer, void* stream);
int hg_timer_stop(void* timer, void* stream, float* ms);
/* hipGraph capture of a device search: capture once, replay many times.  The graph keeps the
 * kernels chosen at capture (the bare-bone search, or the non-bare-bone one when the index held
 * deleted elements then), so hg_graph_launch fails (HG_ERR_RUNTIME) once the GPU index was freed
 * or its delete marks were re-read (hg_gpu_refresh_deleted): capture the search again.
 * Replays of one graph share its launch words: do not overlap them. */
int hg_graph_capture_search(void* gpu_index, const float* d_queries, size_t nq, const int32_t* d_ef_per_query,
                            const int32_t* d_k_per_query, size_t ef, size_t k, size_t k_stride, uint64_t* d_labels,
                            float* d_dists, int32_t* d_counters, void* stop_model, void* stream, void** out_graph);
int hg_graph_launch(void* graph, void* stream);
int hg_graph_free(void* graph);

#ifdef __cplusplus
}
#endif
#endif
