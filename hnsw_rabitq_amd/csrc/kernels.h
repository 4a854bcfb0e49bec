// Host-callable launchers of the gfx950 kernels (implemented in the .hip files).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "common.h"

namespace hg {

// ---- graph search (search.hip) ----
// LDS of one search wave: query, beam, visited hash and the neighbour stage (which the stop-model
// and trace variants also use for the features and the network's activations).
size_t search_lds_bytes(uint32_t dim, uint32_t cap, uint32_t hash_bits);
int search_blocks_per_cu(const SearchParams& p, bool filtered = false);
hipError_t launch_search(const SearchParams& p, int blocks, hipStream_t s);
hipError_t launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s);  // a kernel (graph-safe), not a memset
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s);
hipError_t launch_id_labels(const DevGraph& g, const uint32_t* ids, size_t n, uint64_t* out, hipStream_t s);
// dim / 16 is one of the chunk counts the query kernel is compiled for (what the 8-bit coded rows
// need).
bool search_fixed_chunks(uint32_t dim);
// The 8-bit coded copy of vec ([n][stride] f32) in the DevGraph::vec8 layout; *flag (cleared by
// the caller) gets bit 0 when a coordinate is not an integer in 0..255 with that FP32 value's
// bits (then out is not a lossless copy and must not be used).
hipError_t launch_code_u8(const float* vec, uint32_t n, uint32_t dim, uint32_t stride, uint8_t* out, uint32_t* flag,
                          hipStream_t s);

// ---- multi-vector search (mvsearch.hip) ----
size_t mv_lds_bytes(uint32_t dim, uint32_t cap, uint32_t hash_bits, uint32_t dcap);
int mv_blocks_per_cu(const MvParams& p);
hipError_t launch_mv_search(const MvParams& p, int blocks, hipStream_t s);

// ---- level-0 graph construction (build.hip) ----
constexpr int kBuildUnionCap = 1024;  // list + arrivals re-selected at once by the link kernel
struct BuildArgs {
    DevGraph g;                 // vec, stride, dim, maxM0, space
    uint32_t* links0;           // [n][maxM0], written
    uint32_t s, B;              // batch [s, s + B)
    uint32_t efc, M;
    const uint64_t* cand_ids;   // [B][efc] graph-search candidates (element ids, ~0 = none)
    const float* cand_d;        // [B][efc]
    uint32_t* sel;              // [B][M] selected neighbours (SENT padded)
    // reverse links of the batch, sorted by the element that receives them (all on the device)
    const uint32_t* keys;       // [B*M] receiving element of each link, ascending (then unused slots)
    const uint32_t* incoming;   // [B*M] new element of each link, ascending within a receiver
    const uint32_t* starts;     // [*n_targets] first link of each receiver
    const uint32_t* n_targets;  // number of receivers (device word)
    const uint32_t* n_links;    // number of links (device word)
};
size_t build_select_lds(uint32_t dim, uint32_t efc);
size_t build_link_lds(uint32_t dim);
// link = false: build_select_kernel over the batch; true: build_link_kernel over the receivers
// (launched with B*M workgroups, those beyond *n_targets exit at once)
hipError_t launch_build(const BuildArgs& a, bool link, hipStream_t s);
// Scratch of build_group_links for batches of up to max_links links.
size_t build_group_scratch_bytes(size_t max_links);
// Reverse links of the batch [s, s + B) from sel: pairs (receiver, new element), a stable radix
// sort by receiver, and the first link of each receiver; fills keys / incoming / starts /
// n_targets / n_links of a (all device memory inside `scratch`, no host synchronisation).
// Sets *err to 1 when a selected neighbour is not an earlier element.
hipError_t build_group_links(BuildArgs& a, void* scratch, size_t scratch_bytes, uint32_t* err, hipStream_t s);

// ---- brute force (bruteforce.hip) ----
struct BfData {
    const float* vec;        // [n][dim] f32
    const uint64_t* labels;  // [n]
    const uint16_t* vec_bf16;  // [n][dim] (optional, built on demand)
    const uint8_t* vec_fp8;    // [n][dim] e4m3 (optional)
    const float* fp8_scale;    // [n] per-vector scale (optional)
    const float* norms;        // [n] squared norms in FP32 (L2 GEMM form), optional
    const uint32_t* allow;     // [ceil(n/32)] bitmap of rows a query may return (exact path), optional
    uint32_t n, dim;
    int32_t space;
};

// Exact FP32 search: k smallest (distance, label) pairs per query, rows closest first.
// Needs scratch of bf_exact_scratch_bytes().
size_t bf_exact_scratch_bytes(uint32_t nq, uint32_t k, uint32_t n);
// scratch_cap (bytes, 0 = bf_exact_scratch_bytes) bounds the distance chunk: fewer queries per
// pass when the scratch is smaller (at least one row of n distances).
hipError_t launch_bf_exact(const BfData& d, const float* q, uint32_t nq, uint32_t k, uint64_t* labels, float* dists,
                           void* scratch, hipStream_t s, size_t scratch_cap = 0);

// Reduced-precision candidate pass + FP32 re-rank.
size_t bf_lowp_scratch_bytes(uint32_t nq, uint32_t dim, uint32_t c, uint32_t n);
hipError_t launch_bf_convert(const BfData& d, uint16_t* vec_bf16, uint8_t* vec_fp8, float* fp8_scale, float* norms,
                             hipStream_t s);
// What one reduced-precision search did besides the re-rank.
struct LowpReport {
    uint32_t exact_overflow = 0;     // queries answered by the exact path: candidate list overflowed,
                                     // came up short or had too many ties at the cut
    uint32_t exact_uncertified = 0;  // ... the error bound could not certify the re-ranked list
    uint64_t near_ties = 0, near_tie_disagree = 0;            // FP64 check of adjacent near-tie pairs
    uint64_t boundary_near_ties = 0, boundary_disagree = 0;   // the same at the k-th boundary
    float max_bound = 0.f;           // largest per-query error bound
    bool certified = false;          // the bound was applied
    uint64_t reranked_rows = 0;      // rows whose exact FP32 distance the re-rank computed
};
// Quantisation statistics of the stored rows (3 uint32: float bits of max |x|, max |x~|,
// max |x - x~|), folded in by quant_stats_kernel into out (cleared by the caller).
hipError_t launch_quant_stats(const BfData& d, int precision, uint32_t* out, hipStream_t s);
// xstats (may be null: no certification) = launch_quant_stats's output for these rows;
// tie_rel > 0 runs the FP64 near-tie check in the re-rank.
hipError_t launch_bf_lowp(const BfData& d, int precision, const float* q, uint32_t nq, uint32_t k, uint32_t c,
                          uint64_t* labels, float* dists, void* scratch, hipStream_t s, double* stats,
                          int variant = -1, const uint32_t* xstats = nullptr, double tie_rel = 0.0,
                          LowpReport* rep = nullptr);
constexpr int kGemmVariants = 5;  // HG_GEMM_VARIANTS of the C ABI
void set_gemm_variant(int v);
// FP64 report of final lists: labels are mapped to rows through (sorted_labels, sorted_rows), the
// index's labels ascending and their rows; out_host[5]; scratch >= 64 bytes.
hipError_t launch_fp64_check(const BfData& d, const float* q, uint32_t nq, uint32_t k, const uint64_t* labels,
                             const float* dists, double rel_tol, const uint64_t* sorted_labels,
                             const uint32_t* sorted_rows, double* out_host, void* scratch, hipStream_t s);

// ---- learned stop (mlp.hip) ----
hipError_t launch_mlp_forward(const float* params, int in_dim, int hidden, const float* x, uint32_t n, float* logits,
                              hipStream_t s);
hipError_t launch_mlp_train_step(float* params, float* m, float* v, int step, int in_dim, int hidden, const float* x,
                                 const float* y, uint32_t n, float lr, float b1, float b2, float eps, float* loss_dev,
                                 float* grad_dev, hipStream_t s);

}  // namespace hg
