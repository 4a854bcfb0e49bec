// Brute-force k-NN on gfx950.
//
// Exact path (BruteforceSearch::searchKnn, bruteforce.h:106-125): the reference keeps the k
// lexicographically smallest (distance, label) pairs of a full scan.  Here:
//   1. distance tiles: exact FP32 distances in the reference's order for a chunk of queries
//      against every row, written to a scratch matrix;
//   2. selection: one workgroup per query streams its row of the matrix, keeps every entry
//      below the running k-th (distance, label) key in an LDS buffer, compacts by sorting when
//      the buffer fills, and writes the k smallest closest first.
//
// Reduced-precision path (BF16 / FP8 e4m3 MFMA with FP32 accumulation): the distance GEMM
// produces approximate scores; a per-query threshold keeps about c candidates; their exact
// FP32 distances (reference order) are recomputed and the k smallest (distance, label) pairs
// are selected as above.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"
#include "dist.h"
#include "kernels.h"

namespace hg {

// ------------------------------------------------------------------------------------------
// Exact distance tile, dim % 16 == 0.  32 queries x 32 rows per workgroup of 256 threads.  The
// reference keeps 16 accumulator lanes per (query, row) pair, acc[l] taking coordinates l, l+16,
// l+32, ... in increasing order; here the 16 lanes of a pair sit in 16 threads: thread (group g,
// lane l) keeps lane l of 8 queries x 8 rows (queries 8*(g/4) + i, rows g%4 + 4*j), so one
// coordinate of 8 queries and 8 rows (16 LDS reads) feeds 64 FMAs, and the dimension streams
// through LDS in chunks of 64 coordinates without changing the order of any accumulation.  At
// the end the 16 lanes of each pair are summed in the reference's order through LDS.
// ------------------------------------------------------------------------------------------
constexpr int TQ = 32, TR = 32, TC = 64;
constexpr int TS = 80;  // LDS row stride (floats): rows one apart fall 16 banks apart

template <int SPACE>
__global__ __launch_bounds__(256) void bf_dist_tile_kernel(const float* __restrict__ Q, uint32_t nq,
                                                           const float* __restrict__ X, uint32_t n, uint32_t dim,
                                                           float* __restrict__ out, size_t ld_out) {
    __shared__ __align__(16) float Qs[TQ][TS];
    __shared__ __align__(16) float Xs[TR][TS];
    const int tid = threadIdx.x, l = tid & 15, grp = tid >> 4;
    const int qb = (grp >> 2) * 8, rb = grp & 3;
    // query tiles vary fastest, so the tiles that share a row tile run together and the rows
    // come from HBM about once
    const uint32_t qt = (nq + TQ - 1) / TQ;
    const uint32_t q0 = (blockIdx.x % qt) * TQ, r0 = (blockIdx.x / qt) * TR;
    float acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 8; j++) acc[i][j] = 0.f;

    // this thread's two float4 granules of the query and of the row chunk (32 x 64 floats each),
    // the next chunk's loaded into registers while the current one is used
    static_assert(TQ * (TC / 4) == 512 && TR == TQ, "two granules per thread");
    float4 pq[2], px[2];
    auto gload = [&](uint32_t c0) {
        const uint32_t cw = dim - c0 < (uint32_t)TC ? dim - c0 : (uint32_t)TC;
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int i = tid + 256 * k, rr = i / (TC / 4), cc = (i % (TC / 4)) * 4;
            pq[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            px[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (q0 + rr < nq && (uint32_t)cc < cw)
                pq[k] = *reinterpret_cast<const float4*>(Q + (size_t)(q0 + rr) * dim + c0 + cc);
            if (r0 + rr < n && (uint32_t)cc < cw)
                px[k] = *reinterpret_cast<const float4*>(X + (size_t)(r0 + rr) * dim + c0 + cc);
        }
    };
    gload(0);
    for (uint32_t c0 = 0; c0 < dim; c0 += TC) {
        const uint32_t cw = dim - c0 < (uint32_t)TC ? dim - c0 : (uint32_t)TC;  // multiple of 16
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int i = tid + 256 * k, rr = i / (TC / 4), cc = (i % (TC / 4)) * 4;
            *reinterpret_cast<float4*>(&Qs[rr][cc]) = pq[k];
            *reinterpret_cast<float4*>(&Xs[rr][cc]) = px[k];
        }
        __syncthreads();
        if (c0 + TC < dim) gload(c0 + TC);
        for (uint32_t c = 0; c < cw; c += 16) {
            float qv[8], xv[8];
#pragma unroll
            for (int i = 0; i < 8; i++) qv[i] = Qs[qb + i][c + l];
#pragma unroll
            for (int j = 0; j < 8; j++) xv[j] = Xs[rb + 4 * j][c + l];
#pragma unroll
            for (int i = 0; i < 8; i++)
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    if (SPACE == 0) {
                        const float d = qv[i] - xv[j];
                        acc[i][j] = fma_(d, d, acc[i][j]);
                    } else {
                        acc[i][j] = fma_(qv[i], xv[j], acc[i][j]);
                    }
                }
        }
        __syncthreads();
    }
    // the 16 lanes of each pair, one query index i at a time: lane values to LDS (pair p = 8g + j
    // at p * 17, padded against bank conflicts), then one thread per pair sums them
    float* R = &Qs[0][0];
#pragma unroll
    for (int i = 0; i < 8; i++) {
#pragma unroll
        for (int j = 0; j < 8; j++) R[(grp * 8 + j) * 17 + l] = acc[i][j];
        __syncthreads();
        if (tid < 128) {
            const float* A = R + tid * 17;
            float r;
            if (SPACE == 0) {
                r = A[0];
#pragma unroll
                for (int m = 1; m < 16; m++) r = r + A[m];
            } else {
                float h8[8], h4[4];
#pragma unroll
                for (int m = 0; m < 8; m++) h8[m] = A[m] + A[m + 8];
#pragma unroll
                for (int m = 0; m < 4; m++) h4[m] = h8[m] + h8[m + 4];
                float s0 = h4[0] + h4[2], s1 = h4[1] + h4[3];
                r = 1.0f - (s0 + s1);
            }
            const int g = tid >> 3, j = tid & 7;
            const uint32_t qq = q0 + (g >> 2) * 8 + i, rr = r0 + (g & 3) + 4 * j;
            if (qq < nq && rr < n) out[(size_t)qq * ld_out + rr] = r;
        }
        __syncthreads();
    }
}

// Any dimension: one thread per (query, row) pair, the reference kernel for the dimension class.
__global__ __launch_bounds__(256) void bf_dist_generic_kernel(int space, const float* __restrict__ Q, uint32_t nq,
                                                              const float* __restrict__ X, uint32_t n, uint32_t dim,
                                                              uint32_t stride, float* __restrict__ out,
                                                              size_t ld_out) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t q = blockIdx.y;
    if (r >= n || q >= nq) return;
    out[(size_t)q * ld_out + r] = dist_generic(space, Q + (size_t)q * dim, X + (size_t)r * stride, (int)dim);
}

// ------------------------------------------------------------------------------------------
// Selection: k smallest (distance, label) of one row of `dists` (length n), closest first.
// Optional `rows` maps position -> base row (candidate lists); labels come from base rows.
// ------------------------------------------------------------------------------------------
constexpr int SEL_THREADS = 256;
constexpr int SEL_CAP = 4096;    // LDS candidate buffer
constexpr int SEL_CHUNK = 2048;  // elements scanned between compactions

struct SelEntry {
    float d;
    uint32_t r;
};

__device__ __forceinline__ bool sel_less(const SelEntry& a, const SelEntry& b, const uint64_t* labels) {
    if (a.d < b.d) return true;
    if (b.d < a.d) return false;
    if (a.r == b.r) return false;
    if (a.r == SENT) return false;
    if (b.r == SENT) return true;
    return labels[a.r] < labels[b.r];
}

// Bitonic sort of buf[0, cnt) (padded to a power of two with sentinels) by (distance, label).
template <int NT = SEL_THREADS>
__device__ void sel_sort(SelEntry* buf, int cnt, const uint64_t* labels) {
    int size = 1;
    while (size < cnt) size <<= 1;
    for (int i = cnt + (int)threadIdx.x; i < size; i += NT) buf[i] = SelEntry{INFINITY, SENT};
    __syncthreads();
    for (int k = 2; k <= size; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < size; i += NT) {
                int p = i ^ j;
                if (p > i) {
                    bool up = (i & k) == 0;
                    SelEntry a = buf[i], b = buf[p];
                    bool sw = up ? sel_less(b, a, labels) : sel_less(a, b, labels);
                    if (sw) {
                        buf[i] = b;
                        buf[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(SEL_THREADS) void bf_select_kernel(const float* __restrict__ dists, size_t ld,
                                                                 uint32_t n, const uint32_t* __restrict__ rows,
                                                                 size_t ld_rows, const uint64_t* __restrict__ labels,
                                                                 uint32_t k, uint32_t q_offset,
                                                                 uint64_t* __restrict__ out_l, float* __restrict__ out_d,
                                                                 size_t ld_outp, const uint32_t* __restrict__ allow) {
    __shared__ SelEntry buf[SEL_CAP + SEL_CHUNK];
    __shared__ int cnt;
    __shared__ SelEntry thr;
    const uint32_t q = blockIdx.x;
    const float* row = dists + (size_t)q * ld;
    const uint32_t* rmap = rows ? rows + (size_t)q * ld_rows : nullptr;
    if (threadIdx.x == 0) {
        cnt = 0;
        thr = SelEntry{INFINITY, SENT};
    }
    __syncthreads();
    // this thread's values of a chunk (SEL_CHUNK / SEL_THREADS of them), the next chunk's loaded
    // while the current one is filtered, so the stream keeps two chunks of loads in flight
    constexpr int PER = SEL_CHUNK / SEL_THREADS;
    static_assert(SEL_CHUNK % SEL_THREADS == 0, "whole chunks per thread");
    float dv[PER];
    auto load = [&](uint32_t base) {
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const uint32_t i = base + threadIdx.x + (uint32_t)u * SEL_THREADS;
            dv[u] = i < n ? row[i] : INFINITY;
        }
    };
    load(0);
    for (uint32_t base = 0; base < n; base += SEL_CHUNK) {
        SelEntry t = thr;
        float cur[PER];
#pragma unroll
        for (int u = 0; u < PER; u++) cur[u] = dv[u];
        if (base + SEL_CHUNK < n) load(base + SEL_CHUNK);
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const uint32_t i = base + threadIdx.x + (uint32_t)u * SEL_THREADS;
            if (i >= n) continue;
            uint32_t r = rmap ? rmap[i] : i;
            if (r == SENT) continue;
            // the filter of BruteforceSearch::searchKnn (bruteforce.h:115-117): a row whose label
            // is not allowed never enters the result
            if (allow && ((allow[r >> 5] >> (r & 31u)) & 1u) == 0u) continue;
            SelEntry e{cur[u], r};
            if (sel_less(e, t, labels) || t.r == SENT) {
                int pos = atomicAdd(&cnt, 1);
                buf[pos] = e;
            }
        }
        __syncthreads();
        if (cnt > SEL_CAP - SEL_CHUNK || base + SEL_CHUNK >= n) {
            int c = cnt;
            sel_sort(buf, c, labels);
            int keep = c < (int)k ? c : (int)k;
            if (threadIdx.x == 0) {
                cnt = keep;
                if (keep == (int)k) thr = buf[k - 1];
            }
            __syncthreads();
        }
    }
    int c = cnt;
    uint64_t* ol = out_l + (size_t)(q + q_offset) * ld_outp;
    float* od = out_d + (size_t)(q + q_offset) * ld_outp;
    for (uint32_t i = threadIdx.x; i < k; i += SEL_THREADS) {
        if ((int)i < c) {
            ol[i] = labels[buf[i].r];
            od[i] = buf[i].d;
        } else {
            ol[i] = ~0ull;
            od[i] = INFINITY;
        }
    }
}

size_t bf_exact_scratch_bytes(uint32_t nq, uint32_t k, uint32_t n) {
    (void)k;
    // a chunk of queries x n distances, at most 4 GiB (at least one query row): at 1M rows, 1,024
    // queries per pass, so that the selection has four workgroups per CU
    size_t per_q = (size_t)n * 4;
    size_t qc = std::max<size_t>(1, std::min<size_t>(nq, ((size_t)1 << 32) / std::max<size_t>(per_q, 1)));
    if (qc >= (size_t)TQ) qc = qc / TQ * TQ;  // whole query tiles, within the cap
    return qc * per_q;
}

hipError_t launch_bf_exact(const BfData& d, const float* q, uint32_t nq, uint32_t k, uint64_t* labels, float* dists,
                           void* scratch, hipStream_t s, size_t scratch_cap) {
    size_t scratch_bytes = bf_exact_scratch_bytes(nq, k, d.n);
    uint32_t qc = (uint32_t)(scratch_bytes / ((size_t)d.n * 4));
    if (scratch_cap && scratch_cap < scratch_bytes) {
        // a smaller scratch: as many queries per chunk as fit, whole query tiles when possible
        qc = (uint32_t)std::min<size_t>(qc, scratch_cap / ((size_t)d.n * 4));
        if (qc == 0) return hipErrorInvalidValue;
        if (qc >= (uint32_t)TQ) qc = qc / TQ * TQ;
    }
    float* mat = (float*)scratch;
    for (uint32_t q0 = 0; q0 < nq; q0 += qc) {
        uint32_t cq = nq - q0 < qc ? nq - q0 : qc;
        const float* qp = q + (size_t)q0 * d.dim;
        if (d.dim % 16 == 0) {
            dim3 grid((uint32_t)(((size_t)d.n + TR - 1) / TR * ((cq + TQ - 1) / TQ)));
            if (d.space == 0)
                hipLaunchKernelGGL(bf_dist_tile_kernel<0>, grid, dim3(256), 0, s, qp, cq, d.vec, d.n, d.dim, mat,
                                   (size_t)d.n);
            else
                hipLaunchKernelGGL(bf_dist_tile_kernel<1>, grid, dim3(256), 0, s, qp, cq, d.vec, d.n, d.dim, mat,
                                   (size_t)d.n);
        } else {
            dim3 grid((d.n + 255) / 256, cq);
            hipLaunchKernelGGL(bf_dist_generic_kernel, grid, dim3(256), 0, s, (int)d.space, qp, cq, d.vec, d.n, d.dim,
                               d.dim, mat, (size_t)d.n);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(bf_select_kernel, dim3(cq), dim3(SEL_THREADS), 0, s, mat, (size_t)d.n, d.n,
                           (const uint32_t*)nullptr, (size_t)0, d.labels, k, q0, labels, dists, (size_t)k, d.allow);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ==========================================================================================
// Reduced precision: BF16 / FP8 (OCP e4m3, per-vector scale) MFMA distance GEMM.
// ==========================================================================================
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef int i32x8_t __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// float -> OCP FP8 E4M3 (bias 7, max 448, no infinities), round to nearest even, saturating.
__device__ __forceinline__ uint8_t f32_to_e4m3(float x) {
    uint8_t sign = x < 0.f ? 0x80 : 0x00;
    float a = fabsf(x);
    if (!(a > 0.f)) return sign;  // zero (and NaN mapped to zero)
    if (a >= 448.f) return sign | 0x7E;
    int e;
    float m = frexpf(a, &e);  // a = m * 2^e, m in [0.5, 1)
    int ex = e - 1;           // a = (2m) * 2^ex, 2m in [1, 2)
    if (ex < -6) {
        // subnormal: units of 2^-9
        float q = rintf(ldexpf(a, 9));
        return sign | (uint8_t)q;  // q == 8 encodes the smallest normal
    }
    float mant = rintf((2.f * m - 1.f) * 8.f);
    if (mant >= 8.f) {
        mant = 0.f;
        ex += 1;
    }
    int code = ((ex + 7) << 3) | (int)mant;
    if (code > 0x7E) code = 0x7E;
    return sign | (uint8_t)code;
}

// Row conversion: bf16 copy, or e4m3 copy with scale amax/448 per row; optional squared norms.
__global__ __launch_bounds__(256) void convert_rows_kernel(const float* __restrict__ X, uint32_t n, uint32_t dim,
                                                           uint16_t* __restrict__ out_bf16,
                                                           uint8_t* __restrict__ out_fp8, float* __restrict__ scale) {
    const uint32_t row = blockIdx.x;
    if (row >= n) return;
    const float* x = X + (size_t)row * dim;
    __shared__ float red[256];
    float amax = 0.f;
    for (uint32_t i = threadIdx.x; i < dim; i += 256) amax = fmaxf(amax, fabsf(x[i]));
    red[threadIdx.x] = amax;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    amax = red[0];
    if (out_bf16) {
        for (uint32_t i = threadIdx.x; i < dim; i += 256) out_bf16[(size_t)row * dim + i] = f32_to_bf16_rne(x[i]);
        if (scale && threadIdx.x == 0) scale[row] = 1.0f;
    } else {
        float sc = amax > 0.f ? amax / 448.f : 1.0f;
        float inv = 1.0f / sc;
        for (uint32_t i = threadIdx.x; i < dim; i += 256) out_fp8[(size_t)row * dim + i] = f32_to_e4m3(x[i] * inv);
        if (threadIdx.x == 0) scale[row] = sc;
    }
}

// OCP e4m3 code -> its value (the codes the converter writes: no NaN).
__device__ __forceinline__ double e4m3_value(uint8_t c) {
    const int e = (c >> 3) & 15, m = c & 7;
    const double v = e == 0 ? (double)m * 0x1p-9 : (1.0 + (double)m * 0.125) * ldexp(1.0, e - 7);
    return (c & 0x80) ? -v : v;
}

// Quantisation statistics of rows (FP64, rounded up to FP32): the norm of the FP32 row, the
// norm of its reduced-precision copy (value = code x scale) and the norm of the difference.
// With `each` the three values go to row r of out ([n][3]); otherwise each is folded into
// out[0..2] with an integer max (non-negative floats order as their bits).
__global__ __launch_bounds__(256) void quant_stats_kernel(const float* __restrict__ X, uint32_t n, uint32_t dim,
                                                          const uint16_t* __restrict__ x_bf16,
                                                          const uint8_t* __restrict__ x_fp8,
                                                          const float* __restrict__ scale, int each,
                                                          uint32_t* __restrict__ out) {
    const uint32_t row = blockIdx.x;
    if (row >= n) return;
    const float* x = X + (size_t)row * dim;
    const double sc = x_fp8 ? (double)scale[row] : 1.0;
    double a = 0, b = 0, c = 0;
    for (uint32_t i = threadIdx.x; i < dim; i += 256) {
        const double v = x[i];
        const double t = x_bf16 ? (double)__uint_as_float((uint32_t)x_bf16[(size_t)row * dim + i] << 16)
                                : sc * e4m3_value(x_fp8[(size_t)row * dim + i]);
        a += v * v;
        b += t * t;
        c += (v - t) * (v - t);
    }
    __shared__ double red[3][256];
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    red[2][threadIdx.x] = c;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int j = 0; j < 3; j++) red[j][threadIdx.x] += red[j][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 3) {
        // the FP64 sums are accurate to ~1e-13 relative: a 1e-6 margin and rounding up keep
        // the FP32 value an upper bound
        const float v = __double2float_ru(sqrt(red[threadIdx.x][0]) * (1.0 + 1e-6));
        if (each)
            out[(size_t)row * 3 + threadIdx.x] = __float_as_uint(v);
        else
            atomicMax(&out[threadIdx.x], __float_as_uint(v));
    }
}

// Per-query bound E on |approximate distance - exact FP32 distance| over every row, from the
// query's statistics (qs[q] = ‖q‖, ‖q~‖, ‖q - q~‖) and the rows' maxima (xs = max ‖x‖, max
// ‖x~‖, max ‖x - x~‖), with u = 2^-24 and g = D u / (1 - D u) (the bound of an FP32 sum of D
// terms in any order):
//   dot part   |fl(q~.x~) - q.x| <= (2g + 3u)(1 + 3u) ‖q~‖‖x~‖ + ‖q‖ ‖x - x~‖ + ‖q - q~‖ ‖x~‖
//              (q.x - q~.x~ = q.(x - x~) + (q - q~).x~; the MFMA sum, whose rounding mode is not
//              assumed, so 2u per addition, and the two scale products)
//   IP         E = dot + g ‖q‖‖x‖ + u (1 + ‖q‖‖x‖) + u (1 + ‖q~‖‖x~‖)   (1 - s, both sides)
//   L2         E = 2 dot + g (‖q‖² + ‖x‖²) [the FP32 norms] + 2u (‖q‖² + ‖x‖² + ‖q~‖‖x~‖)
//              [the two roundings of a = (nq + nx) - 2 s] + g' (‖q‖ + ‖x‖)² [the exact form's
//              differences and sum, g' for D + 2 terms]
// A row that the candidate pass left out has approximate distance >= A (the cut); its exact
// distance is then >= A - E, so when A - E exceeds the k-th exact distance found, no left-out
// row can be among the k smallest and the list equals the exact search's.
__global__ void lowp_bound_kernel(uint32_t nq, uint32_t dim, int space, const uint32_t* __restrict__ qs,
                                  const uint32_t* __restrict__ xs, float* __restrict__ bound) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const double u = 0x1p-24;
    const double g = (double)dim * u / (1.0 - (double)dim * u);
    const double g2 = (double)(dim + 2) * u / (1.0 - (double)(dim + 2) * u);
    const double qn = __uint_as_float(qs[(size_t)q * 3]), qt = __uint_as_float(qs[(size_t)q * 3 + 1]),
                 qe = __uint_as_float(qs[(size_t)q * 3 + 2]);
    const double xn = __uint_as_float(xs[0]), xt = __uint_as_float(xs[1]), xe = __uint_as_float(xs[2]);
    const double dot = (2 * g + 3 * u) * (1 + 3 * u) * qt * xt + qn * xe + qe * xt;
    double e;
    if (space == 0)
        e = 2 * dot + g * (qn * qn + xn * xn) + 2 * u * (qn * qn + xn * xn + qt * xt) + g2 * (qn + xn) * (qn + xn);
    else
        e = dot + g * qn * xn + u * (1 + qn * xn) + u * (1 + qt * xt);
    bound[q] = __double2float_ru(e * 1.001);
}

// Tile order shared by the GEMM kernels.  Workgroup b runs on XCD b % 8.  The query tiles are
// split into 8 groups, one per XCD, so a group's query rows stay in that XCD's L2; inside a
// group the row tile varies slowest and every XCD walks the row tiles in the same order, so a
// row tile is read from HBM about once and shared through the memory-side cache.
struct TileOrder {
    uint32_t tiles_m, tiles_n, G, groups;
    __host__ __device__ static TileOrder make(uint32_t tm, uint32_t tn) {
        TileOrder o;
        o.tiles_m = tm;
        o.tiles_n = tn;
        o.G = (tm + 7) / 8;
        o.groups = (tm + o.G - 1) / o.G;
        return o;
    }
    __host__ __device__ uint32_t grid() const { return 8u * ((groups + 7) / 8) * G * tiles_n; }
    // false when this workgroup has no tile
    __device__ __forceinline__ bool map(uint32_t b, uint32_t& tm, uint32_t& tn) const {
        const uint32_t x = b % 8, local = b / 8;
        const uint32_t per_group = G * tiles_n;
        const uint32_t g = x + 8 * (local / per_group);
        const uint32_t rem = local % per_group;
        if (g >= groups) return false;
        tn = rem / G;
        tm = g * G + rem % G;
        return tm < tiles_m;
    }
};

// GEMM tile: 128 queries x 128 rows per workgroup of 4 waves (2 x 2 waves of 64 x 64), each
// wave 2 x 2 MFMA 32x32 blocks.  LDS holds 128-byte k-slices of both operands (64 bf16 or 128
// fp8 per row), double buffered.  MFMA operand layout (gfx950): lane l supplies row l % 32 and
// the contiguous k run starting at (l / 32) * (k per lane); the 32x32 result holds
// (row (i/4)*8 + (l/32)*4 + i%4, column l % 32) in register i of lane l.
constexpr int GM = 128, GN = 128, GKB = 128, GPAD = 16;  // GKB = bytes of k per stage

struct LowpGemmArgs {
    const uint8_t* A;        // queries, [nq][dim] low precision
    const float* a_scale;    // [nq]
    const uint8_t* B;        // rows, [n][dim] low precision
    const float* b_scale;    // [n]
    const float* b_norm;     // [n] squared norms (L2), FP32
    const float* a_norm;     // [nq] squared norms (L2)
    uint32_t nq, n, dim;     // dim in elements
    int32_t space;
    // sample mode: write every approximate distance
    float* out_mat;
    size_t ld_out;
    // filter mode: keep rows whose approximate distance is below tau[q]
    const float* tau;
    uint32_t* cand_rows;     // [nq][cap]
    float* cand_d;           // [nq][cap]
    uint32_t* cand_cnt;      // [nq]
    uint32_t cap;
    uint32_t tiles_m, tiles_n;
};

// Per-row and per-column values of one workgroup tile, staged in LDS for the epilogue.
template <int BM, int BN, int WN>
struct __align__(16) EpiLds {
    float tau[BM], a_norm[BM], a_scale[BM];
    float bsc[BN], bn[BN];
    uint32_t cnt[BM];      // passes of each row in the tile (LDS adds, cleared before the main loop)
    uint32_t base[BM];     // start of the row's reserved candidate slots
    uint32_t ovf;          // a wave's pass list outgrew its LDS space: the tile takes pass 2
};

// The tile's per-row and per-column epilogue inputs: thread t loads those of row m0 + t and of
// column n0 + t.  Issued before the first k-slices (so they are the oldest loads in flight and
// the main loop's first wait covers them) and written to LDS after that wait, so their latency
// hides behind the pipeline fill and the registers are free again for the loop.
struct EpiVals {
    float tau, a_norm, a_scale, bsc, bn;
};

template <int PREC, bool FILTER, int BM, int BN, int NT>
__device__ __forceinline__ EpiVals epi_load(const LowpGemmArgs& g, uint32_t m0, uint32_t n0, int tid) {
    static_assert(BM <= NT && BN <= NT, "one row and one column of the tile per thread");
    EpiVals v;
    const uint32_t row = m0 + tid, col = n0 + tid;
    const bool rok = tid < BM && row < g.nq, cok = tid < BN && col < g.n;
    v.tau = (FILTER && rok) ? g.tau[row] : -INFINITY;
    v.a_norm = (g.space == 0 && rok) ? g.a_norm[row] : 0.f;
    v.a_scale = (PREC == 2 && rok) ? g.a_scale[row] : 1.f;
    v.bsc = (PREC == 2 && cok) ? g.b_scale[col] : 1.0f;
    v.bn = (g.space == 0 && cok) ? g.b_norm[col] : 0.f;
    return v;
}

template <int BM, int BN, int WN>
__device__ __forceinline__ void epi_store(const EpiVals& v, int tid, EpiLds<BM, BN, WN>& e) {
    if (tid < BM) {
        e.tau[tid] = v.tau;
        e.a_norm[tid] = v.a_norm;
        e.a_scale[tid] = v.a_scale;
    }
    if (tid < BN) {
        e.bsc[tid] = v.bsc;
        e.bn[tid] = v.bn;
    }
    if (tid < BM) e.cnt[tid] = 0u;
    if (tid == 0) e.ovf = 0u;
}

// Four per-row values of one 32x32 block held by a half-wave: registers 4q..4q+3 of the
// accumulator belong to rows lb + 8q + 0..3, where lb includes the half-wave's offset of 4.
template <typename T>
__device__ __forceinline__ void rows4(const T* a, int lb, int q, T (&v)[4]) {
    typedef T T4 __attribute__((ext_vector_type(4)));
    const T4 t = *reinterpret_cast<const T4*>(a + lb + q * 8);
    v[0] = t.x;
    v[1] = t.y;
    v[2] = t.z;
    v[3] = t.w;
}

// Epilogue of the GEMM variants.  The workgroup tile is rows [m0, m0 + 64*TM) x columns
// [n0, n0 + 32*TN*WN); wave (wm, wn) holds TM x TN 32x32 blocks in acc, so each of its 32*TM
// rows sits in one (block row i, register r, half-wave h) and the WN waves wn share a row.
// Sample mode writes every approximate distance.  Filter mode appends the columns with
// d < tau[row] to the row's candidate list with one device-scope atomic per (row, workgroup):
// pass 1 counts each row's passes (LDS adds) and lists each wave's passes in LDS, one thread
// per row reserves the row's total, and each wave then writes its list out (a list that
// outgrows its LDS space sends the tile to pass 2, which recomputes the passes instead).
// Per-row values are read from LDS four rows at a time (one 16-byte read) before each group
// of four pairs.  NT threads; two waves along the rows.  SPACE is fixed at compile time so
// that the unrolled loops carry no branch.
template <int PREC, bool FILTER, int TM, int TN, int WN, int NT, int SPACE>
__device__ __forceinline__ void gemm_epilogue_s(const LowpGemmArgs& g, f32x16_t (&acc)[TM][TN], uint32_t m0,
                                                uint32_t n0, int wm, int wn, int lane, int tid,
                                                EpiLds<64 * TM, 32 * TN * WN, WN>& e, uint2* cbuf, uint32_t capw) {
    constexpr int BM = 64 * TM;
    __syncthreads();
    uint32_t col[TN];
    bool col_ok[TN];
    float bsc[TN], bn[TN];
#pragma unroll
    for (int j = 0; j < TN; j++) {
        const int c = wn * 32 * TN + j * 32 + (lane & 31);
        col[j] = n0 + c;
        col_ok[j] = col[j] < g.n;
        bsc[j] = PREC == 2 ? e.bsc[c] : 1.0f;
        bn[j] = SPACE == 0 ? e.bn[c] : 0.f;
    }
    auto dist = [&](int i, int j, int r, float an, float as) {
        float sv = acc[i][j][r];
        // opaque to the compiler, so that a pass recomputes its distances instead of keeping
        // another pass's 16 * TM * TN values live across the barriers
        __asm__ volatile("" : "+v"(sv));
        if (PREC == 2) sv = sv * as * bsc[j];
        if (SPACE == 0) return an + bn[j] - 2.0f * sv;
        return 1.0f - sv;
    };
    const int half = lane >> 5;
    // per-row values of rows lb + 8q + 0..3 for this half-wave
    float an[4] = {}, as[4] = {};
    auto load_rows = [&](int lb, int q) {
        if (SPACE == 0) rows4(e.a_norm, lb, q, an);
        if (PREC == 2) rows4(e.a_scale, lb, q, as);
    };
    if (!FILTER) {
#pragma unroll
        for (int i = 0; i < TM; i++) {
            const int lb = wm * 32 * TM + i * 32 + half * 4;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                load_rows(lb, q);
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
                    const uint32_t row = m0 + lb + q * 8 + rr;
#pragma unroll
                    for (int j = 0; j < TN; j++)
                        if (col_ok[j] && row < g.nq)
                            g.out_mat[(size_t)row * g.ld_out + col[j]] = dist(i, j, q * 4 + rr, an[rr], as[rr]);
                }
            }
        }
        return;
    }
    // pass 1: each passing lane adds one to its row's count in LDS (no return value is used,
    // so the adds do not wait); `hit` marks the (i, r) pairs with any pass (wave-uniform, in
    // scalar registers).  Each wave also lists its passes in its own part of the idle operand
    // LDS as (row << 16 | column within the tile, distance bits); wcnt is wave-uniform.
    uint64_t hit = 0, colm[TN];
#pragma unroll
    for (int j = 0; j < TN; j++) colm[j] = __builtin_amdgcn_ballot_w64(col_ok[j]);
    uint2* wl = cbuf + (size_t)(tid >> 6) * capw;
    uint32_t wcnt = 0;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
#pragma unroll
    for (int i = 0; i < TM; i++) {
        const int lb = wm * 32 * TM + i * 32 + half * 4;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            float tq[4];
            rows4(e.tau, lb, q, tq);
            load_rows(lb, q);
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                const int r = q * 4 + rr;
                const uint32_t lr = (uint32_t)(lb + q * 8 + rr);
#pragma unroll
                for (int j = 0; j < TN; j++) {
                    const float dv = dist(i, j, r, an[rr], as[rr]);
                    const uint64_t m = __builtin_amdgcn_ballot_w64(dv < tq[rr]) & colm[j];
                    if (m) {  // wave-uniform
                        hit |= 1ull << (i * 16 + r);
                        if ((m >> lane) & 1ull) {
                            atomicAdd(&e.cnt[lr], 1u);
                            const uint32_t pos = wcnt + (uint32_t)__popcll(m & lt);
                            if (pos < capw)
                                wl[pos] = make_uint2((lr << 16) | (uint32_t)(wn * 32 * TN + j * 32 + (lane & 31)),
                                                     __float_as_uint(dv));
                        }
                        wcnt += (uint32_t)__popcll(m);
                    }
                }
            }
        }
    }
    if (wcnt > capw && lane == 0) e.ovf = 1u;
    __syncthreads();
    // one reservation per row
    for (int r = tid; r < BM; r += NT) {
        const uint32_t c = e.cnt[r];
        if (c) e.base[r] = atomicAdd(&g.cand_cnt[m0 + r], c);
    }
    __syncthreads();
    // a row's slots are handed out by an LDS counter that starts at the row's reserved base (the
    // order of candidates inside a row's list carries no meaning)
    if (e.ovf == 0u) {
        // every wave writes its list out, 64 candidates per store
        for (uint32_t idx = lane; idx < wcnt; idx += 64) {
            const uint2 v = wl[idx];
            const uint32_t lr = v.x >> 16;
            const uint32_t slot = atomicAdd(&e.base[lr], 1u);
            if (slot < g.cap) {
                const size_t o = (size_t)(m0 + lr) * g.cap + slot;
                g.cand_rows[o] = n0 + (v.x & 0xFFFFu);
                g.cand_d[o] = __uint_as_float(v.y);
            }
        }
        return;
    }
    // pass 2 (a list overflowed): recompute the marked pairs and write their passes
#pragma unroll
    for (int i = 0; i < TM; i++) {
        const int lb = wm * 32 * TM + i * 32 + half * 4;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (!((hit >> (i * 16 + q * 4)) & 0xFull)) continue;  // no pass in this group: scalar skip
            float tq[4];
            rows4(e.tau, lb, q, tq);
            load_rows(lb, q);
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                const int r = q * 4 + rr;
                if (!((hit >> (i * 16 + r)) & 1ull)) continue;  // no pass in this pair: scalar skip
                const uint32_t lr = (uint32_t)(lb + q * 8 + rr);
#pragma unroll
                for (int j = 0; j < TN; j++) {
                    const float dv = dist(i, j, r, an[rr], as[rr]);
                    if (col_ok[j] && dv < tq[rr]) {
                        const uint32_t slot = atomicAdd(&e.base[lr], 1u);
                        if (slot < g.cap) {
                            g.cand_rows[(size_t)(m0 + lr) * g.cap + slot] = col[j];
                            g.cand_d[(size_t)(m0 + lr) * g.cap + slot] = dv;
                        }
                    }
                }
            }
        }
    }
}

template <int PREC, bool FILTER, int TM, int TN, int WN, int NT>
__device__ __forceinline__ void gemm_epilogue(const LowpGemmArgs& g, f32x16_t (&acc)[TM][TN], uint32_t m0,
                                              uint32_t n0, int wm, int wn, int lane, int tid,
                                              EpiLds<64 * TM, 32 * TN * WN, WN>& e, uint2* cbuf,
                                              uint32_t capw) {
    if (g.space == 0)
        gemm_epilogue_s<PREC, FILTER, TM, TN, WN, NT, 0>(g, acc, m0, n0, wm, wn, lane, tid, e, cbuf, capw);
    else
        gemm_epilogue_s<PREC, FILTER, TM, TN, WN, NT, 1>(g, acc, m0, n0, wm, wn, lane, tid, e, cbuf, capw);
}

template <int PREC, bool FILTER>
__global__ __launch_bounds__(256, 2) void lowp_gemm_kernel(LowpGemmArgs g) {
    __shared__ __align__(16) uint8_t As[2][GM][GKB + GPAD];
    __shared__ __align__(16) uint8_t Bs[2][GN][GKB + GPAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    uint32_t tm, tn;
    if (!TileOrder::make(g.tiles_m, g.tiles_n).map(blockIdx.x, tm, tn)) return;
    const uint32_t m0 = tm * GM, n0 = tn * GN;
    __shared__ EpiLds<64 * 2, 128, 2> epi;
    epi_store(epi_load<PREC, FILTER, 128, 128, 256>(g, m0, n0, tid), tid, epi);
    const size_t row_bytes = (size_t)g.dim * (PREC == 1 ? 2 : 1);

    f32x16_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.f;

    // each thread moves 4 x 16 bytes of A and of B per stage
    uint4 ra[4], rb[4];
    auto gload = [&](size_t kb) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            int idx = tid + 256 * i;
            int row = idx >> 3, ch = idx & 7;
            uint32_t qa = m0 + row, qb = n0 + row;
            // an FP8 row of dim % 128 == 64 ends halfway through the last stage: the chunks past
            // its end are zero (a zero code adds nothing to the dot product)
            const bool in_row = kb + ch * 16 < row_bytes;
            ra[i] = qa < g.nq && in_row ? *reinterpret_cast<const uint4*>(g.A + qa * row_bytes + kb + ch * 16)
                                        : make_uint4(0, 0, 0, 0);
            rb[i] = qb < g.n && in_row ? *reinterpret_cast<const uint4*>(g.B + qb * row_bytes + kb + ch * 16)
                                       : make_uint4(0, 0, 0, 0);
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            int idx = tid + 256 * i;
            int row = idx >> 3, ch = idx & 7;
            *reinterpret_cast<uint4*>(&As[buf][row][ch * 16]) = ra[i];
            *reinterpret_cast<uint4*>(&Bs[buf][row][ch * 16]) = rb[i];
        }
    };
    const size_t kbytes = row_bytes;
    gload(0);
    lstore(0);
    __syncthreads();
    int buf = 0;
    for (size_t kb = 0; kb < kbytes; kb += GKB) {
        const bool more = kb + GKB < kbytes;
        if (more) gload(kb + GKB);
        // MFMAs on the current stage
        if (PREC == 1) {
#pragma unroll
            for (int ks = 0; ks < GKB; ks += 32) {  // 16 bf16 = 32 bytes per MFMA k-step
                bf16x8_t fa[2], fb[2];
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    const uint8_t* pa = &As[buf][wm * 64 + i * 32 + (lane & 31)][ks + (lane >> 5) * 16];
                    const uint8_t* pb = &Bs[buf][wn * 64 + i * 32 + (lane & 31)][ks + (lane >> 5) * 16];
                    fa[i] = *reinterpret_cast<const bf16x8_t*>(pa);
                    fb[i] = *reinterpret_cast<const bf16x8_t*>(pb);
                }
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int j = 0; j < 2; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int ks = 0; ks < GKB; ks += 64) {  // 64 fp8 per MFMA k-step
                i32x8_t fa[2], fb[2];
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    const uint8_t* pa = &As[buf][wm * 64 + i * 32 + (lane & 31)][ks + (lane >> 5) * 32];
                    const uint8_t* pb = &Bs[buf][wn * 64 + i * 32 + (lane & 31)][ks + (lane >> 5) * 32];
                    fa[i] = *reinterpret_cast<const i32x8_t*>(pa);
                    fb[i] = *reinterpret_cast<const i32x8_t*>(pb);
                }
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int j = 0; j < 2; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fa[i], fb[j], acc[i][j], 0, 0, 0,
                                                                                    127, 0, 127);
            }
        }
        if (more) {
            lstore(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    }

    // the A operand's LDS (36 KiB) holds the pass lists once the main loop is done
    gemm_epilogue<PREC, FILTER, 2, 2, 2, 256>(g, acc, m0, n0, wm, wn, lane, tid, epi, reinterpret_cast<uint2*>(&As[0][0][0]),
                                              (uint32_t)(sizeof(As) / 8 / 4));
}

// Larger tiles: each wave owns (32*TM) x (32*TN) outputs (TM x TN MFMA blocks), the workgroup
// 2 x 2 waves.  One LDS stage (k-slice of 128 bytes per row for both operands); the next slice
// is prefetched into registers during the MFMAs and stored after a barrier, so a workgroup
// needs (64*TM + 64*TN) * 144 bytes of LDS and two of them fit on a CU.
template <int PREC, bool FILTER, int TM, int TN>
__global__ __launch_bounds__(256, 2) void lowp_gemm_wide_kernel(LowpGemmArgs g) {
    constexpr int BM = 64 * TM, BN = 64 * TN;
    constexpr int LA = BM * 8 / 256, LB = BN * 8 / 256;  // 16-byte chunks per thread
    __shared__ __align__(16) uint8_t As[BM][GKB + GPAD];
    __shared__ __align__(16) uint8_t Bs[BN][GKB + GPAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    uint32_t tm, tn;
    if (!TileOrder::make((g.nq + BM - 1) / BM, (g.n + BN - 1) / BN).map(blockIdx.x, tm, tn)) return;
    const uint32_t m0 = tm * BM, n0 = tn * BN;
    __shared__ EpiLds<BM, BN, 2> epi;
    epi_store(epi_load<PREC, FILTER, BM, BN, 256>(g, m0, n0, tid), tid, epi);
    const size_t row_bytes = (size_t)g.dim * (PREC == 1 ? 2 : 1);

    f32x16_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.f;

    uint4 ra[LA], rb[LB];
    auto gload = [&](size_t kb) {
#pragma unroll
        for (int i = 0; i < LA; i++) {
            int idx = tid + 256 * i;
            int row = idx >> 3, ch = idx & 7;
            uint32_t qa = m0 + row;
            // chunks past the end of an FP8 row of dim % 128 == 64 are zero
            ra[i] = qa < g.nq && kb + ch * 16 < row_bytes
                        ? *reinterpret_cast<const uint4*>(g.A + qa * row_bytes + kb + ch * 16)
                        : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < LB; i++) {
            int idx = tid + 256 * i;
            int row = idx >> 3, ch = idx & 7;
            uint32_t qb = n0 + row;
            rb[i] = qb < g.n && kb + ch * 16 < row_bytes
                        ? *reinterpret_cast<const uint4*>(g.B + qb * row_bytes + kb + ch * 16)
                        : make_uint4(0, 0, 0, 0);
        }
    };
    auto lstore = [&]() {
#pragma unroll
        for (int i = 0; i < LA; i++) {
            int idx = tid + 256 * i;
            *reinterpret_cast<uint4*>(&As[idx >> 3][(idx & 7) * 16]) = ra[i];
        }
#pragma unroll
        for (int i = 0; i < LB; i++) {
            int idx = tid + 256 * i;
            *reinterpret_cast<uint4*>(&Bs[idx >> 3][(idx & 7) * 16]) = rb[i];
        }
    };
    const size_t kbytes = row_bytes;
    gload(0);
    lstore();
    __syncthreads();
    for (size_t kb = 0; kb < kbytes; kb += GKB) {
        const bool more = kb + GKB < kbytes;
        if (more) gload(kb + GKB);
        if (PREC == 1) {
#pragma unroll
            for (int ks = 0; ks < GKB; ks += 32) {
                bf16x8_t fa[TM], fb[TN];
#pragma unroll
                for (int i = 0; i < TM; i++)
                    fa[i] = *reinterpret_cast<const bf16x8_t*>(&As[wm * 32 * TM + i * 32 + (lane & 31)][ks + (lane >> 5) * 16]);
#pragma unroll
                for (int j = 0; j < TN; j++)
                    fb[j] = *reinterpret_cast<const bf16x8_t*>(&Bs[wn * 32 * TN + j * 32 + (lane & 31)][ks + (lane >> 5) * 16]);
#pragma unroll
                for (int i = 0; i < TM; i++)
#pragma unroll
                    for (int j = 0; j < TN; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int ks = 0; ks < GKB; ks += 64) {
                i32x8_t fa[TM], fb[TN];
#pragma unroll
                for (int i = 0; i < TM; i++)
                    fa[i] = *reinterpret_cast<const i32x8_t*>(&As[wm * 32 * TM + i * 32 + (lane & 31)][ks + (lane >> 5) * 32]);
#pragma unroll
                for (int j = 0; j < TN; j++)
                    fb[j] = *reinterpret_cast<const i32x8_t*>(&Bs[wn * 32 * TN + j * 32 + (lane & 31)][ks + (lane >> 5) * 32]);
#pragma unroll
                for (int i = 0; i < TM; i++)
#pragma unroll
                    for (int j = 0; j < TN; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fa[i], fb[j], acc[i][j], 0, 0, 0,
                                                                                    127, 0, 127);
            }
        }
        if (more) {
            __syncthreads();
            lstore();
            __syncthreads();
        }
    }

    // the A operand's LDS holds the pass lists once the main loop is done
    gemm_epilogue<PREC, FILTER, TM, TN, 2, 256>(g, acc, m0, n0, wm, wn, lane, tid, epi, reinterpret_cast<uint2*>(&As[0][0]),
                                                (uint32_t)(sizeof(As) / 8 / 4));
}

// Direct-to-LDS pipeline: rows are fetched with global_load_lds_dwordx4 (64 lanes x 16 bytes
// = 1 KiB of LDS per wave instruction, no register staging) into a ring of three 64-byte
// k-slices, two slices ahead of the MFMAs.  Inside a 64-byte row the four 16-byte chunks are
// XOR-swizzled by (row >> 1) & 3 so that the 32 rows read by one MFMA operand spread over all
// LDS banks.  Workgroup tile 256 x 128 (4 waves of 4 x 2 MFMA 32x32 blocks).
constexpr int LKB = 64, LNST = 3;

// 16-byte chunk permutation of a 64-byte row: the rows read together by 8 (or 16) lanes of a
// ds_read_b128 cover every bank once with 32 (or 64) four-byte banks.
__device__ __forceinline__ int swzf(int row) { return ((row >> 1) & 3) ^ ((row >> 3) & 1); }
__device__ __forceinline__ int swz(int row, int chunk) { return row * LKB + ((chunk ^ swzf(row)) << 4); }

// Three slices per operand as distinct LDS objects, so the compiler's wait-count pass can tell
// which in-flight direct-to-LDS loads an LDS read depends on (only the slice being read).
template <int TM, int TN>
struct LdsRing {
    __attribute__((address_space(3))) uint8_t* a[LNST];
    __attribute__((address_space(3))) uint8_t* b[LNST];
};

template <int PREC, int TM, int TN>  // TM: A blocks of 16 rows per wave, TN: B blocks per wave
__device__ __forceinline__ void gemm_lds_issue(const LowpGemmArgs& g, __attribute__((address_space(3))) uint8_t* dA,
                                               __attribute__((address_space(3))) uint8_t* dB, int st, uint32_t m0,
                                               uint32_t n0, size_t row_bytes, int wave, int lane) {
    const int lrow = lane >> 2, lslot = lane & 3;
    const size_t kb = (size_t)st * LKB;
#pragma unroll
    for (int i = 0; i < TM; i++) {  // A: TM blocks of 16 rows per wave
        const int blk = wave * TM + i;
        const int row = blk * 16 + lrow;
        uint32_t q = m0 + row;
        q = q < g.nq ? q : g.nq - 1;
        const int c = lslot ^ swzf(row);
        __builtin_amdgcn_global_load_lds((const void*)(g.A + (size_t)q * row_bytes + kb + c * 16), dA + blk * 1024, 16,
                                         0, 0);
    }
#pragma unroll
    for (int i = 0; i < TN; i++) {  // B: TN blocks of 16 rows per wave
        const int blk = wave * TN + i;
        const int row = blk * 16 + lrow;
        uint32_t r = n0 + row;
        r = r < g.n ? r : g.n - 1;
        const int c = lslot ^ swzf(row);
        __builtin_amdgcn_global_load_lds((const void*)(g.B + (size_t)r * row_bytes + kb + c * 16), dB + blk * 1024, 16,
                                         0, 0);
    }
}

typedef int i32x4_t __attribute__((ext_vector_type(4)));

// Six 16-byte LDS reads completed inside one asm block (reads and their wait together, outputs
// early-clobbered), so the compiler's wait-count pass does not make the read wait for the
// direct-to-LDS loads still in flight for later slices; the caller orders the slice itself
// (vmcnt wait + barrier).
__device__ __forceinline__ void lds_read6(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, uint32_t a4, uint32_t a5,
                                          i32x4_t& r0, i32x4_t& r1, i32x4_t& r2, i32x4_t& r3, i32x4_t& r4,
                                          i32x4_t& r5) {
    asm volatile(
        "ds_read_b128 %0, %6\n\t"
        "ds_read_b128 %1, %7\n\t"
        "ds_read_b128 %2, %8\n\t"
        "ds_read_b128 %3, %9\n\t"
        "ds_read_b128 %4, %10\n\t"
        "ds_read_b128 %5, %11\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(r0), "=&v"(r1), "=&v"(r2), "=&v"(r3), "=&v"(r4), "=&v"(r5)
        : "v"(a0), "v"(a1), "v"(a2), "v"(a3), "v"(a4), "v"(a5)
        : "memory");
}

__device__ __forceinline__ void lds_read8(const uint32_t (&a)[8], i32x4_t (&r)[8]) {
    asm volatile(
        "ds_read_b128 %0, %8\n\t"
        "ds_read_b128 %1, %9\n\t"
        "ds_read_b128 %2, %10\n\t"
        "ds_read_b128 %3, %11\n\t"
        "ds_read_b128 %4, %12\n\t"
        "ds_read_b128 %5, %13\n\t"
        "ds_read_b128 %6, %14\n\t"
        "ds_read_b128 %7, %15\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&v"(r[3]), "=&v"(r[4]), "=&v"(r[5]), "=&v"(r[6]), "=&v"(r[7])
        : "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(a[4]), "v"(a[5]), "v"(a[6]), "v"(a[7])
        : "memory");
}

// 4 x 4 blocks per wave (128 x 128 outputs): eight operand reads feed sixteen MFMAs.
template <int PREC>
__device__ __forceinline__ void gemm_lds_mfma44(uint32_t A, uint32_t B, f32x16_t (&acc)[4][4], int wm, int wn,
                                                int lane) {
    const int ra = wm * 128 + (lane & 31), rb = wn * 128 + (lane & 31);
    if (PREC == 1) {
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            const int c = ks * 2 + (lane >> 5);
            const uint32_t ad[8] = {A + swz(ra, c), A + swz(ra + 32, c), A + swz(ra + 64, c), A + swz(ra + 96, c),
                                    B + swz(rb, c), B + swz(rb + 32, c), B + swz(rb + 64, c), B + swz(rb + 96, c)};
            i32x4_t f[8];
            lds_read8(ad, f);
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, f[i]),
                                                                        __builtin_bit_cast(bf16x8_t, f[4 + j]),
                                                                        acc[i][j], 0, 0, 0);
        }
    } else {
        const int c = (lane >> 5) * 2;
        const uint32_t al[8] = {A + swz(ra, c), A + swz(ra + 32, c), A + swz(ra + 64, c), A + swz(ra + 96, c),
                                B + swz(rb, c), B + swz(rb + 32, c), B + swz(rb + 64, c), B + swz(rb + 96, c)};
        const uint32_t ah[8] = {A + swz(ra, c + 1), A + swz(ra + 32, c + 1), A + swz(ra + 64, c + 1),
                                A + swz(ra + 96, c + 1), B + swz(rb, c + 1), B + swz(rb + 32, c + 1),
                                B + swz(rb + 64, c + 1), B + swz(rb + 96, c + 1)};
        i32x4_t lo[8], hi[8];
        lds_read8(al, lo);
        lds_read8(ah, hi);
        i32x8_t op[8];
#pragma unroll
        for (int t = 0; t < 8; t++)
            op[t] = i32x8_t{lo[t].x, lo[t].y, lo[t].z, lo[t].w, hi[t].x, hi[t].y, hi[t].z, hi[t].w};
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(op[i], op[4 + j], acc[i][j], 0, 0, 0,
                                                                            127, 0, 127);
    }
}

template <int PREC, int TM, int TN>
__device__ __forceinline__ void gemm_lds_mfma(uint32_t A, uint32_t B, f32x16_t (&acc)[TM][TN], int wm, int wn,
                                              int lane) {
    static_assert(TM == 4 && TN == 2, "fragment reads are written for 4 x 2 blocks per wave");
    if (PREC == 1) {
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {  // two 32x32x16 steps per 64-byte slice
            const int c = ks * 2 + (lane >> 5);
            const int ra = wm * 128 + (lane & 31), rb = wn * 64 + (lane & 31);
            i32x4_t f0, f1, f2, f3, g0, g1;
            lds_read6(A + swz(ra, c), A + swz(ra + 32, c), A + swz(ra + 64, c), A + swz(ra + 96, c), B + swz(rb, c),
                      B + swz(rb + 32, c), f0, f1, f2, f3, g0, g1);
            const bf16x8_t fa[4] = {__builtin_bit_cast(bf16x8_t, f0), __builtin_bit_cast(bf16x8_t, f1),
                                    __builtin_bit_cast(bf16x8_t, f2), __builtin_bit_cast(bf16x8_t, f3)};
            const bf16x8_t fb[2] = {__builtin_bit_cast(bf16x8_t, g0), __builtin_bit_cast(bf16x8_t, g1)};
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
    } else {
        const int c = (lane >> 5) * 2;
        const int ra = wm * 128 + (lane & 31), rb = wn * 64 + (lane & 31);
        i32x4_t l0, l1, l2, l3, m0, m1, h0, h1, h2, h3, n0, n1;
        lds_read6(A + swz(ra, c), A + swz(ra + 32, c), A + swz(ra + 64, c), A + swz(ra + 96, c), B + swz(rb, c),
                  B + swz(rb + 32, c), l0, l1, l2, l3, m0, m1);
        lds_read6(A + swz(ra, c + 1), A + swz(ra + 32, c + 1), A + swz(ra + 64, c + 1), A + swz(ra + 96, c + 1),
                  B + swz(rb, c + 1), B + swz(rb + 32, c + 1), h0, h1, h2, h3, n0, n1);
        const i32x8_t fa[4] = {i32x8_t{l0.x, l0.y, l0.z, l0.w, h0.x, h0.y, h0.z, h0.w},
                               i32x8_t{l1.x, l1.y, l1.z, l1.w, h1.x, h1.y, h1.z, h1.w},
                               i32x8_t{l2.x, l2.y, l2.z, l2.w, h2.x, h2.y, h2.z, h2.w},
                               i32x8_t{l3.x, l3.y, l3.z, l3.w, h3.x, h3.y, h3.z, h3.w}};
        const i32x8_t fb[2] = {i32x8_t{m0.x, m0.y, m0.z, m0.w, n0.x, n0.y, n0.z, n0.w},
                               i32x8_t{m1.x, m1.y, m1.z, m1.w, n1.x, n1.y, n1.z, n1.w}};
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++)
                acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fa[i], fb[j], acc[i][j], 0, 0, 0, 127, 0,
                                                                            127);
    }
}

// WNW waves along the columns, TN 32-column blocks per wave (WNW 2, TN 2: 256 x 128 tile, 256
// threads, two workgroups per CU; WNW 4, TN 2: 256 x 256, 512 threads; WNW 2, TN 4: 256 x 256,
// 256 threads with 4 x 4 blocks per wave, accumulators beyond the VGPRs held in AGPRs, a third
// fewer operand reads per MFMA).
template <int PREC, bool FILTER, int WNW, int TN>
__global__ __launch_bounds__(128 * WNW, (WNW == 2 && TN == 2) ? 2 : 1) void lowp_gemm_lds_kernel(LowpGemmArgs g) {
    constexpr int NT = 128 * WNW, NW = 2 * WNW;
    constexpr int TM = 4, BM = 64 * TM, BN = 32 * TN * WNW;
    constexpr int APW = BM / 16 / NW, BPW = BN / 16 / NW;  // 16-row blocks per wave and slice
    __shared__ __align__(1024) uint8_t sA0[BM * LKB], sA1[BM * LKB], sA2[BM * LKB];
    __shared__ __align__(1024) uint8_t sB0[BN * LKB], sB1[BN * LKB], sB2[BN * LKB];
    typedef __attribute__((address_space(3))) uint8_t* lp;
    lp A0 = (lp)sA0, A1 = (lp)sA1, A2 = (lp)sA2, B0 = (lp)sB0, B1 = (lp)sB1, B2 = (lp)sB2;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WNW, wn = wave % WNW;
    uint32_t tm, tn;
    if (!TileOrder::make((g.nq + BM - 1) / BM, (g.n + BN - 1) / BN).map(blockIdx.x, tm, tn)) return;
    const uint32_t m0 = tm * BM, n0 = tn * BN;
    __shared__ EpiLds<BM, BN, WNW> epi;
    const EpiVals ev = epi_load<PREC, FILTER, BM, BN, NT>(g, m0, n0, tid);
    const size_t row_bytes = (size_t)g.dim * (PREC == 1 ? 2 : 1);
    const int nst = (int)(row_bytes / LKB);
    constexpr int PER_STAGE = APW + BPW;  // direct-to-LDS loads per wave per slice

    f32x16_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.f;

    gemm_lds_issue<PREC, APW, BPW>(g, A0, B0, 0, m0, n0, row_bytes, wave, lane);
    if (nst > 1) gemm_lds_issue<PREC, APW, BPW>(g, A1, B1, 1, m0, n0, row_bytes, wave, lane);
    // one step: wait for slice st (the loads of slice st+1 may stay in flight), barrier, issue
    // slice st+2 into the ring slot freed by slice st-1, MFMAs on slice st
    auto step = [&](int st, lp rA, lp rB, lp nA, lp nB) {
        if (st + 1 < nst)
            __builtin_amdgcn_s_waitcnt(0x0F70 | PER_STAGE);  // vmcnt <= loads of one slice
        else
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt == 0
        __builtin_amdgcn_s_barrier();
        if (st + 2 < nst) gemm_lds_issue<PREC, APW, BPW>(g, nA, nB, st + 2, m0, n0, row_bytes, wave, lane);
        if constexpr (TN == 4)
            gemm_lds_mfma44<PREC>((uint32_t)(uintptr_t)rA, (uint32_t)(uintptr_t)rB, acc, wm, wn, lane);
        else
            gemm_lds_mfma<PREC, TM, TN>((uint32_t)(uintptr_t)rA, (uint32_t)(uintptr_t)rB, acc, wm, wn, lane);
    };
    // nst >= 1 (dim is a multiple of 64): the first step, then the epilogue inputs (complete by its
    // wait) to LDS, then the ring continues from slot 1
    step(0, A0, B0, A2, B2);
    epi_store(ev, tid, epi);
    for (int st = 1; st < nst; st += 3) {
        step(st, A1, B1, A0, B0);
        if (st + 1 < nst) step(st + 1, A2, B2, A1, B1);
        if (st + 2 < nst) step(st + 2, A0, B0, A2, B2);
    }

    // the first A slice's LDS holds the pass lists once the main loop is done
    gemm_epilogue<PREC, FILTER, TM, TN, WNW, NT>(g, acc, m0, n0, wm, wn, lane, tid, epi, reinterpret_cast<uint2*>(sA0),
                                                 (uint32_t)(sizeof(sA0) / 8 / NW));
}

// Copy every `step`-th row of a low-precision matrix (and its scale / norm) to a compact sample.
__global__ void gather_rows_kernel(const uint8_t* __restrict__ src, size_t row_bytes, uint32_t step, uint32_t s_rows,
                                   uint8_t* __restrict__ dst, const float* sc, float* dsc, const float* nm,
                                   float* dnm) {
    uint32_t r = blockIdx.x;
    if (r >= s_rows) return;
    const uint8_t* a = src + (size_t)r * step * row_bytes;
    uint8_t* b = dst + (size_t)r * row_bytes;
    for (size_t i = threadIdx.x * 16; i < row_bytes; i += blockDim.x * 16)
        *reinterpret_cast<uint4*>(b + i) = *reinterpret_cast<const uint4*>(a + i);
    if (threadIdx.x == 0) {
        if (sc) dsc[r] = sc[(size_t)r * step];
        if (nm) dnm[r] = nm[(size_t)r * step];
    }
}

// Squared norms in FP32 (L2 form of the approximate distance).
__global__ void row_norms_kernel(const float* __restrict__ X, uint32_t n, uint32_t dim, float* __restrict__ out) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float* x = X + (size_t)r * dim;
    float s = 0.f;
    for (uint32_t i = 0; i < dim; i++) s = fma_(x[i], x[i], s);
    out[r] = s;
}

__device__ __forceinline__ uint32_t f2key(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key2f(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(u);
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// tau[q] = the rank-th smallest value of row q of the sample matrix (radix select, 4 x 8 bits),
// nudged up by one ulp so that exactly the sample values below it and at it pass `d < tau`.
constexpr uint32_t KTH_MAX = 16384;  // sample rows (the sample step keeps it at most this)

__global__ __launch_bounds__(256) void kth_kernel(const float* __restrict__ mat, size_t ld, uint32_t s,
                                                  uint32_t rank, float* __restrict__ tau) {
    __shared__ uint32_t hist[256];
    __shared__ uint32_t prefix, want;
    const float* row = mat + (size_t)blockIdx.x * ld;
    if (threadIdx.x == 0) {
        prefix = 0;
        want = rank;  // 1-based rank among the remaining candidates
    }
    // the row's keys, read once: element threadIdx.x + 256 u in register u
    constexpr int PER = KTH_MAX / 256;
    uint32_t keys[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const uint32_t i = threadIdx.x + 256u * (uint32_t)u;
        keys[u] = i < s ? f2key(row[i]) : 0u;
    }
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    for (int shift = 24; shift >= 0; shift -= 8) {
        hist[threadIdx.x] = 0;
        __syncthreads();
        const uint32_t pre = prefix;
        const uint32_t mask_hi = shift == 24 ? 0u : (0xFFFFFFFFu << (shift + 8));
        // most keys share a few buckets (the high bytes of similar distances): lanes whose bucket
        // is the wave's first active lane's add together with one atomic, the others one each
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const uint32_t i = threadIdx.x + 256u * (uint32_t)u;
            const uint32_t k = keys[u];
            const bool in = i < s && (k & mask_hi) == pre;
            const uint32_t bkt = (k >> shift) & 255u;
            const uint64_t act = __ballot(in);
            if (act) {
                const uint32_t lead = (uint32_t)__builtin_amdgcn_readlane((int)bkt, __ffsll((unsigned long long)act) - 1);
                const uint64_t same = __ballot(in && bkt == lead);
                if (in && bkt == lead) {
                    if (lane == (uint32_t)(__ffsll((unsigned long long)same) - 1))
                        atomicAdd(&hist[bkt], (uint32_t)__popcll(same));
                } else if (in) {
                    atomicAdd(&hist[bkt], 1u);
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t w = want, acc = 0, bsel = 255;
            for (uint32_t bkt = 0; bkt < 256; bkt++) {
                if (acc + hist[bkt] >= w) {
                    bsel = bkt;
                    break;
                }
                acc += hist[bkt];
            }
            want = w - acc;
            prefix = pre | (bsel << shift);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float v = key2f(prefix);
        tau[blockIdx.x] = nextafterf(v, INFINITY);
    }
}

// Per query: take the candidates of the filter pass (at most cap), keep the c best by
// approximate distance, recompute their exact FP32 distances in the reference order, and write
// the k smallest (distance, label) closest first.  A query whose candidate count overflowed the
// buffer is flagged for the exact path.
constexpr int RR_THREADS = 512;
constexpr int RR_CAP = 16384;  // candidate slots per query
constexpr int RR_TIES = 2048;  // candidates tied at the c-th approximate distance one query can resolve
constexpr int RR_PER = 4096 / RR_THREADS;  // chosen rows per thread (c <= 4096)

// Key of the rank-th smallest approximate distance among buf[0, m) (radix select over the LDS
// entries, 4 x 8 bits); every thread returns it.
template <int NT>
__device__ uint32_t lds_select_key(const SelEntry* buf, uint32_t m, uint32_t rank, uint32_t* hist, uint32_t& prefix,
                                   uint32_t& want) {
    if (threadIdx.x == 0) {
        prefix = 0;
        want = rank;
    }
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int b = threadIdx.x; b < 256; b += NT) hist[b] = 0;
        __syncthreads();
        const uint32_t pre = prefix;
        const uint32_t mask_hi = shift == 24 ? 0u : (0xFFFFFFFFu << (shift + 8));
        for (uint32_t i = threadIdx.x; i < m; i += NT) {
            const uint32_t key = f2key(buf[i].d);
            if ((key & mask_hi) == pre) atomicAdd(&hist[(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t w = want, acc = 0, bsel = 255;
            for (uint32_t bkt = 0; bkt < 256; bkt++) {
                if (acc + hist[bkt] >= w) {
                    bsel = bkt;
                    break;
                }
                acc += hist[bkt];
            }
            want = w - acc;
            prefix = pre | (bsel << shift);
        }
        __syncthreads();
    }
    return prefix;
}

// Re-rank of one query per workgroup.  The c candidates with the smallest approximate distance
// (ties by label) are chosen by a radix select that reads the candidate list from global memory,
// so LDS holds only the c chosen rows; their exact FP32 distances (reference lane order) are
// computed four lanes per row and the k smallest by (distance, label) are written.  A query
// whose candidate list overflowed, came up short, or has more than RR_TIES ties at the cut is
// flagged (1) for the exact path, and so is (2) a query whose list the error bound cannot
// certify: the rows left out have approximate distance >= A (the c-th candidate's, or tau when
// every candidate was taken), so when A - bound[q] <= the k-th exact distance a left-out row
// might belong to the k smallest.  tie_ctr (optional): FP64 check of the near-tie decisions --
// adjacent pairs of the re-ranked order, up to the pair at the k-th boundary, whose FP32
// distances differ by at most rel_tol (relative) are recomputed in FP64; counters [0] near-tie
// pairs, [1] pairs whose FP64 order differs, [2] and [3] the same for the boundary pair.
//
// Bound pruning (with a bound): a chosen row whose approximate distance exceeds a1 + 2 bound[q],
// a1 the (k+1)-th smallest approximate distance among the chosen rows, has exact distance
// > a1 + bound[q] >= the (k+1)-th smallest exact distance among them (the k + 1 rows of smallest
// approximate distance all lie at or below a1 + bound[q]), so it cannot take any of the first
// k + 1 places of the exact order: only the other rows are read in FP32, and the k results, the
// certification and the near-tie pairs (up to the one at the k-th boundary) are those of the
// full re-rank.  rr_ctr[0] += rows re-ranked.
template <int SPACE>
__global__ __launch_bounds__(RR_THREADS) void rerank_kernel(const float* __restrict__ Q, const float* __restrict__ X,
                                                            const uint64_t* __restrict__ labels, uint32_t dim,
                                                            const uint32_t* __restrict__ cand_rows,
                                                            const float* __restrict__ cand_d,
                                                            const uint32_t* __restrict__ cand_cnt, uint32_t cap,
                                                            uint32_t c, uint32_t k, uint64_t* __restrict__ out_l,
                                                            float* __restrict__ out_d, size_t ld_out,
                                                            uint32_t* __restrict__ overflow, uint32_t n_rows,
                                                            const float* __restrict__ tau,
                                                            const float* __restrict__ bound,
                                                            unsigned long long* __restrict__ tie_ctr, double rel_tol,
                                                            unsigned long long* __restrict__ rr_ctr) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t cpow = 2;  // >= 2 keeps the query copy 16-byte aligned
    while (cpow < c) cpow <<= 1;
    // the tie labels are resolved before the first candidate is compacted, so both share the
    // first region (max(cpow, RR_TIES) 8-byte slots); the query follows
    SelEntry* buf = reinterpret_cast<SelEntry*>(smem);    // cpow entries
    uint64_t* ties = reinterpret_cast<uint64_t*>(smem);   // RR_TIES labels
    const uint32_t slots = cpow > (uint32_t)RR_TIES ? cpow : (uint32_t)RR_TIES;
    float* qs = reinterpret_cast<float*>(buf + slots);     // dim floats
    __shared__ uint32_t hist[256];
    __shared__ uint32_t prefix, want, nsel, ntie, tfill, bad, cut_on, npr;
    __shared__ uint64_t cut_label;
    const uint32_t q = blockIdx.x;
    const uint32_t cnt = cand_cnt[q];
    if (cnt > cap || cnt < k) {
        if (threadIdx.x == 0) overflow[q] = 1;
        return;
    }
    const float* cdq = cand_d + (size_t)q * cap;
    const uint32_t* crq = cand_rows + (size_t)q * cap;
    for (uint32_t i = threadIdx.x; i < dim; i += RR_THREADS) qs[i] = Q[(size_t)q * dim + i];
    const bool all = cnt <= c;
    const uint32_t m = all ? cnt : c;
    // smallest approximate distance a left-out row can have (every row taken: none is left out)
    float acut = all ? (cnt >= n_rows ? INFINITY : tau[q]) : 0.f;
    if (threadIdx.x == 0) {
        prefix = 0;
        want = c;
        nsel = 0;
        ntie = 0;
        tfill = 0;
        cut_label = 0;
        cut_on = 0;
        bad = 0;
        npr = 0;
    }
    __syncthreads();
    if (!all) {
        // key of the c-th smallest approximate distance; `want` ends as the number of entries
        // with that key inside the first c
        for (int shift = 24; shift >= 0; shift -= 8) {
            for (int b = threadIdx.x; b < 256; b += RR_THREADS) hist[b] = 0;
            __syncthreads();
            const uint32_t pre = prefix;
            const uint32_t mask_hi = shift == 24 ? 0u : (0xFFFFFFFFu << (shift + 8));
            for (uint32_t i = threadIdx.x; i < cnt; i += RR_THREADS) {
                const uint32_t key = f2key(cdq[i]);
                if ((key & mask_hi) == pre) atomicAdd(&hist[(key >> shift) & 255u], 1u);
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t w = want, acc = 0, bsel = 255;
                for (uint32_t bkt = 0; bkt < 256; bkt++) {
                    if (acc + hist[bkt] >= w) {
                        bsel = bkt;
                        break;
                    }
                    acc += hist[bkt];
                }
                want = w - acc;
                prefix = pre | (bsel << shift);
                // last digit: hist[bsel] entries carry the cut key
                if (shift == 0 && hist[bsel] > want) ntie = hist[bsel];
            }
            __syncthreads();
        }
        const uint32_t kc = prefix;
        const uint32_t nt = ntie;
        acut = key2f(kc);
        if (nt > 0) {
            // more entries carry the cut key than are taken: keep the `want` smallest labels
            // (the order the exact lists use among equal distances)
            if (nt > (uint32_t)RR_TIES) {
                if (threadIdx.x == 0) overflow[q] = 1;
                return;
            }
            for (uint32_t i = threadIdx.x; i < cnt; i += RR_THREADS)
                if (f2key(cdq[i]) == kc) ties[atomicAdd(&tfill, 1u)] = labels[crq[i]];
            __syncthreads();
            uint32_t tp = 1;
            while (tp < nt) tp <<= 1;
            for (uint32_t i = nt + threadIdx.x; i < tp; i += RR_THREADS) ties[i] = ~0ull;
            __syncthreads();
            for (uint32_t kk = 2; kk <= tp; kk <<= 1)
                for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
                    for (uint32_t i = threadIdx.x; i < tp; i += RR_THREADS) {
                        const uint32_t pp = i ^ j;
                        if (pp > i) {
                            const uint64_t a = ties[i], b = ties[pp];
                            if (((i & kk) == 0) ? a > b : a < b) {
                                ties[i] = b;
                                ties[pp] = a;
                            }
                        }
                    }
                    __syncthreads();
                }
            if (threadIdx.x == 0) {
                cut_label = ties[want - 1];
                cut_on = 1;
            }
            __syncthreads();
        }
        // compaction of the chosen rows
        const uint64_t cl = cut_label;
        const bool con = cut_on != 0;
        for (uint32_t i0 = 0; i0 < cnt; i0 += RR_THREADS) {
            const uint32_t i = i0 + threadIdx.x;
            bool take = false;
            uint32_t r = 0;
            float a = 0.f;
            if (i < cnt) {
                a = cdq[i];
                const uint32_t key = f2key(a);
                r = crq[i];
                take = key < kc || (key == kc && (!con || labels[r] <= cl));
            }
            const uint64_t bm = __ballot(take);
            uint32_t base = 0;
            const int lane = threadIdx.x & 63;
            if (lane == 0 && bm) base = atomicAdd(&nsel, (uint32_t)__popcll(bm));
            base = (uint32_t)__shfl((int)base, 0);
            if (take) {
                const uint32_t slot = base + (uint32_t)__popcll(bm & ((lane ? ~0ull >> (64 - lane) : 0ull)));
                if (slot < m)
                    buf[slot] = SelEntry{a, r};  // approximate distance until the exact pass
                else
                    bad = 1;
            }
        }
        __syncthreads();
        if (nsel != m || bad) {  // cannot happen when the select is right; the exact path answers
            if (threadIdx.x == 0) overflow[q] = 1;
            return;
        }
    } else {
        for (uint32_t i = threadIdx.x; i < cnt; i += RR_THREADS) buf[i] = SelEntry{cdq[i], crq[i]};
        __syncthreads();
    }
    // bound pruning: the rows that can take one of the first k + 1 places, compacted in place
    // (each thread holds its strided share of the chosen rows in registers across a barrier)
    uint32_t mw = m;
    if (bound && m > k + 1) {
        const double a1 = key2f(lds_select_key<RR_THREADS>(buf, m, k + 1, hist, prefix, want));
        const double e2 = 2.0 * (double)bound[q];
        // (the double sum is exact to 2^-53; the margin keeps the test on the safe side)
        const double t = a1 + e2 + (fabs(a1) + e2) * 0x1p-40;
        SelEntry mine[RR_PER];
        bool kp[RR_PER];
#pragma unroll
        for (int u = 0; u < RR_PER; u++) {
            const uint32_t i = threadIdx.x + (uint32_t)u * RR_THREADS;
            kp[u] = i < m && (double)buf[i].d <= t;
            mine[u] = kp[u] ? buf[i] : SelEntry{0.f, 0u};
        }
        __syncthreads();
        const int ln = threadIdx.x & 63;
#pragma unroll
        for (int u = 0; u < RR_PER; u++) {
            const uint64_t bm = __ballot(kp[u]);
            uint32_t base = 0;
            if (ln == 0 && bm) base = atomicAdd(&npr, (uint32_t)__popcll(bm));
            base = (uint32_t)__shfl((int)base, 0);
            if (kp[u]) buf[base + (uint32_t)__popcll(bm & ((ln ? ~0ull >> (64 - ln) : 0ull)))] = mine[u];
        }
        __syncthreads();
        mw = npr;
    }
    if (threadIdx.x == 0) atomicAdd(&rr_ctr[0], (unsigned long long)mw);
    // exact distances, four lanes per row (reference order)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int grp = lane >> 2, sub = lane & 3;
    const int nch = (int)(dim >> 4);
    constexpr int NW = RR_THREADS / 64;
    for (uint32_t base = wave * 16; base < mw; base += NW * 16) {
        const uint32_t idx = base + grp;
        const bool act = idx < mw;
        const uint32_t r = act ? buf[idx].r : 0u;
        float res;
        dist_group16<SPACE, 0>(X + (size_t)r * dim, qs, nch, act, sub, res, lane);
        if (act && sub == 0) buf[idx].d = res;
    }
    __syncthreads();
    sel_sort<RR_THREADS>(buf, (int)mw, labels);
    // certification (uniform over the workgroup: every thread reads the same values)
    if (bound && !((double)acut - (double)bound[q] > (double)buf[k - 1].d)) {
        if (threadIdx.x == 0) overflow[q] = 2;
        return;
    }
    if (tie_ctr) {
        for (uint32_t i = threadIdx.x; i < k && i + 1 < mw; i += RR_THREADS) {
            const double a = buf[i].d, b = buf[i + 1].d;
            if (b - a > rel_tol * fabs(a)) continue;
            const float* xa = X + (size_t)buf[i].r * dim;
            const float* xb = X + (size_t)buf[i + 1].r * dim;
            double sa = 0, sb = 0;
            for (uint32_t j = 0; j < dim; j++) {
                const double qj = qs[j];
                if (SPACE == 0) {
                    const double ta = qj - (double)xa[j], tb = qj - (double)xb[j];
                    sa = fma(ta, ta, sa);
                    sb = fma(tb, tb, sb);
                } else {
                    sa = fma(qj, (double)xa[j], sa);
                    sb = fma(qj, (double)xb[j], sb);
                }
            }
            if (SPACE != 0) {
                sa = 1.0 - sa;
                sb = 1.0 - sb;
            }
            // the FP32 order puts a before b; FP64 disagrees when it puts b first
            const bool dis = sa > sb || (sa == sb && labels[buf[i].r] > labels[buf[i + 1].r]);
            atomicAdd(&tie_ctr[0], 1ull);
            if (dis) atomicAdd(&tie_ctr[1], 1ull);
            if (i == k - 1) {
                atomicAdd(&tie_ctr[2], 1ull);
                if (dis) atomicAdd(&tie_ctr[3], 1ull);
            }
        }
    }
    uint64_t* ol = out_l + (size_t)q * ld_out;
    float* od = out_d + (size_t)q * ld_out;
    for (uint32_t i = threadIdx.x; i < k; i += RR_THREADS) {
        ol[i] = labels[buf[i].r];
        od[i] = buf[i].d;
    }
}

// Sample step of the threshold pass: at most 16,384 sample rows.
static inline uint32_t sample_step(uint32_t n) { return n > 16384 ? (n + 16383) / 16384 : 1; }

size_t bf_lowp_scratch_bytes(uint32_t nq, uint32_t dim, uint32_t c, uint32_t n) {
    (void)c;
    const uint32_t step = sample_step(n);
    const uint32_t s = (n + step - 1) / step;
    size_t b = 0;
    b += (size_t)nq * dim * 2;        // queries in low precision (bf16: 2 bytes)
    b += (size_t)nq * 4 * 4;          // a_scale, a_norm, tau, bound
    b += (size_t)nq * 12;             // query quantisation statistics
    b += (size_t)s * dim * 2 + (size_t)s * 8;  // sample rows + scale + norm
    // sample matrix; afterwards the exact path's distance chunk for the queries it answers
    b += std::max<size_t>((size_t)nq * s, n) * 4;
    b += (size_t)nq * RR_CAP * 8;     // candidates; afterwards the exact path's queries and rows
    b += (size_t)nq * 8;              // counts + flags
    b += (size_t)n * 4;               // row norms
    b += 64;                          // FP64 near-tie counters
    return b + 16 * 256;
}

hipError_t launch_bf_convert(const BfData& d, uint16_t* vec_bf16, uint8_t* vec_fp8, float* fp8_scale, float* norms,
                             hipStream_t s) {
    (void)norms;
    if (d.n == 0) return hipSuccess;
    hipLaunchKernelGGL(convert_rows_kernel, dim3(d.n), dim3(256), 0, s, d.vec, d.n, d.dim, vec_bf16, vec_fp8,
                       fp8_scale);
    return hipGetLastError();
}

hipError_t launch_quant_stats(const BfData& d, int precision, uint32_t* out, hipStream_t s) {
    if (d.n == 0) return hipSuccess;
    hipLaunchKernelGGL(quant_stats_kernel, dim3(d.n), dim3(256), 0, s, d.vec, d.n, d.dim,
                       precision == 1 ? d.vec_bf16 : nullptr, precision == 2 ? d.vec_fp8 : nullptr, d.fp8_scale, 0,
                       out);
    return hipGetLastError();
}

static inline size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// GEMM variant: 0 = 128x128 workgroup tiles (2x2 blocks per wave, double-buffered LDS),
// 1 = 256x128 workgroup tiles (4x2 blocks per wave, register-prefetched single LDS stage),
// 2 = 256x128 workgroup tiles fed by direct-to-LDS loads through a 3-slice ring, 3 = 256x256
// tiles with eight waves on the same ring, 4 = 256x256 tiles with four waves of 4x4 blocks.
// The process default applies when a caller passes variant < 0.
static int g_gemm_variant = 2;
void set_gemm_variant(int v) { g_gemm_variant = v; }

static void launch_gemm(LowpGemmArgs ga, int precision, bool filter, int variant, hipStream_t s) {
    const int g_gemm_variant = variant >= 0 ? variant : hg::g_gemm_variant;
    if (g_gemm_variant == 0) {
        ga.tiles_m = (ga.nq + GM - 1) / GM;
        ga.tiles_n = (ga.n + GN - 1) / GN;
        uint32_t grid = TileOrder::make(ga.tiles_m, ga.tiles_n).grid();
        if (precision == 1) {
            if (filter) hipLaunchKernelGGL((lowp_gemm_kernel<1, true>), dim3(grid), dim3(256), 0, s, ga);
            else hipLaunchKernelGGL((lowp_gemm_kernel<1, false>), dim3(grid), dim3(256), 0, s, ga);
        } else {
            if (filter) hipLaunchKernelGGL((lowp_gemm_kernel<2, true>), dim3(grid), dim3(256), 0, s, ga);
            else hipLaunchKernelGGL((lowp_gemm_kernel<2, false>), dim3(grid), dim3(256), 0, s, ga);
        }
        return;
    }
    if (g_gemm_variant == 4) {
        uint32_t grid = TileOrder::make((ga.nq + 255) / 256, (ga.n + 255) / 256).grid();
        if (precision == 1) {
            if (filter) hipLaunchKernelGGL((lowp_gemm_lds_kernel<1, true, 2, 4>), dim3(grid), dim3(256), 0, s, ga);
            else hipLaunchKernelGGL((lowp_gemm_lds_kernel<1, false, 2, 4>), dim3(grid), dim3(256), 0, s, ga);
        } else {
            if (filter) hipLaunchKernelGGL((lowp_gemm_lds_kernel<2, true, 2, 4>), dim3(grid), dim3(256), 0, s, ga);
            else hipLaunchKernelGGL((lowp_gemm_lds_kernel<2, false, 2, 4>), dim3(grid), dim3(256), 0, s, ga);
        }
        return;
    }
    if (g_gemm_variant == 3) {
        uint32_t grid = TileOrder::make((ga.nq + 255) / 256, (ga.n + 255) / 256).grid();
        if (precision == 1) {
            if (filter) hipLaunchKernelGGL((lowp_gemm_lds_kernel<1, true, 4, 2>), dim3(grid), dim3(512), 0, s, ga);
            else hipLaunchKernelGGL((lowp_gemm_lds_kernel<1, false, 4, 2>), dim3(grid), dim3(512), 0, s, ga);
        } else {
            if (filter) hipLaunchKernelGGL((lowp_gemm_lds_kernel<2, true, 4, 2>), dim3(grid), dim3(512), 0, s, ga);
            else hipLaunchKernelGGL((lowp_gemm_lds_kernel<2, false, 4, 2>), dim3(grid), dim3(512), 0, s, ga);
        }
        return;
    }
    constexpr int TM = 4, TN = 2;
    uint32_t tm = (ga.nq + 64 * TM - 1) / (64 * TM), tn = (ga.n + 64 * TN - 1) / (64 * TN);
    uint32_t grid = TileOrder::make(tm, tn).grid();
    if (g_gemm_variant == 2) {
        if (precision == 1) {
            if (filter) hipLaunchKernelGGL((lowp_gemm_lds_kernel<1, true, 2, 2>), dim3(grid), dim3(256), 0, s, ga);
            else hipLaunchKernelGGL((lowp_gemm_lds_kernel<1, false, 2, 2>), dim3(grid), dim3(256), 0, s, ga);
        } else {
            if (filter) hipLaunchKernelGGL((lowp_gemm_lds_kernel<2, true, 2, 2>), dim3(grid), dim3(256), 0, s, ga);
            else hipLaunchKernelGGL((lowp_gemm_lds_kernel<2, false, 2, 2>), dim3(grid), dim3(256), 0, s, ga);
        }
        return;
    }
    if (precision == 1) {
        if (filter) hipLaunchKernelGGL((lowp_gemm_wide_kernel<1, true, TM, TN>), dim3(grid), dim3(256), 0, s, ga);
        else hipLaunchKernelGGL((lowp_gemm_wide_kernel<1, false, TM, TN>), dim3(grid), dim3(256), 0, s, ga);
    } else {
        if (filter) hipLaunchKernelGGL((lowp_gemm_wide_kernel<2, true, TM, TN>), dim3(grid), dim3(256), 0, s, ga);
        else hipLaunchKernelGGL((lowp_gemm_wide_kernel<2, false, TM, TN>), dim3(grid), dim3(256), 0, s, ga);
    }
}

// Rows r of the exact path's output go back to row idx[r] of the caller's lists.
__global__ void gather_queries_kernel(const float* __restrict__ q, uint32_t dim, const uint32_t* __restrict__ idx,
                                      uint32_t nf, float* __restrict__ out) {
    const uint32_t r = blockIdx.x;
    if (r >= nf) return;
    for (uint32_t i = threadIdx.x; i < dim; i += blockDim.x) out[(size_t)r * dim + i] = q[(size_t)idx[r] * dim + i];
}

__global__ void scatter_rows_kernel(const uint64_t* __restrict__ sl, const float* __restrict__ sd, uint32_t k,
                                    const uint32_t* __restrict__ idx, uint32_t nf, uint64_t* __restrict__ labels,
                                    float* __restrict__ dists) {
    const uint32_t r = blockIdx.x;
    if (r >= nf) return;
    for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
        labels[(size_t)idx[r] * k + i] = sl[(size_t)r * k + i];
        dists[(size_t)idx[r] * k + i] = sd[(size_t)r * k + i];
    }
}

namespace {
struct Events {
    hipEvent_t ev[4] = {};
    Events() {
        for (auto& e : ev) (void)hipEventCreate(&e);
    }
    ~Events() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};
}  // namespace

hipError_t launch_bf_lowp(const BfData& d, int precision, const float* q, uint32_t nq, uint32_t k, uint32_t c,
                          uint64_t* labels, float* dists, void* scratch, hipStream_t s, double* stats, int variant,
                          const uint32_t* xstats, double tie_rel, LowpReport* rep) {
    const uint32_t n = d.n, dim = d.dim;
    const size_t elem = precision == 1 ? 2 : 1;
    const uint32_t step = sample_step(n);
    const uint32_t srows = (n + step - 1) / step;
    if (srows > KTH_MAX) return hipErrorInvalidValue;  // the sample step keeps srows <= 16384
    char* p = (char*)scratch;
    auto take = [&](size_t bytes) {
        char* r = p;
        p += align_up(bytes);
        return r;
    };
    uint8_t* qlp = (uint8_t*)take((size_t)nq * dim * elem);
    float* qsc = (float*)take((size_t)nq * 4);
    float* qnm = (float*)take((size_t)nq * 4);
    float* tau = (float*)take((size_t)nq * 4);
    float* bound = (float*)take((size_t)nq * 4);
    uint32_t* qst = (uint32_t*)take((size_t)nq * 12);
    uint8_t* slp = (uint8_t*)take((size_t)srows * dim * elem);
    float* ssc = (float*)take((size_t)srows * 4);
    float* snm = (float*)take((size_t)srows * 4);
    const size_t smat_bytes = std::max<size_t>((size_t)nq * srows, n) * 4;
    float* smat = (float*)take(smat_bytes);
    uint32_t* crow = (uint32_t*)take((size_t)nq * RR_CAP * 4);
    float* cd = (float*)take((size_t)nq * RR_CAP * 4);
    uint32_t* ccnt = (uint32_t*)take((size_t)nq * 8);  // counts, then flags
    uint32_t* ovf = ccnt + nq;
    float* xnm = (float*)take((size_t)n * 4);
    unsigned long long* tie_ctr = (unsigned long long*)take(64);
    const uint8_t* xlp = precision == 1 ? (const uint8_t*)d.vec_bf16 : d.vec_fp8;
    Events ev;
    (void)hipEventRecord(ev.ev[0], s);

    // queries to low precision; norms for the L2 form; the error bound of each query
    hipLaunchKernelGGL(convert_rows_kernel, dim3(nq), dim3(256), 0, s, q, nq, dim,
                       precision == 1 ? (uint16_t*)qlp : nullptr, precision == 2 ? qlp : nullptr, qsc);
    if (d.space == 0) {
        hipLaunchKernelGGL(row_norms_kernel, dim3((nq + 255) / 256), dim3(256), 0, s, q, nq, dim, qnm);
        hipLaunchKernelGGL(row_norms_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d.vec, n, dim, xnm);
    }
    if (xstats) {
        hipLaunchKernelGGL(quant_stats_kernel, dim3(nq), dim3(256), 0, s, q, nq, dim,
                           precision == 1 ? (const uint16_t*)qlp : nullptr, precision == 2 ? (const uint8_t*)qlp : nullptr,
                           (const float*)qsc, 1, qst);
        hipLaunchKernelGGL(lowp_bound_kernel, dim3((nq + 255) / 256), dim3(256), 0, s, nq, dim, (int)d.space,
                           (const uint32_t*)qst, xstats, bound);
    }
    // sample pass -> per-query threshold
    hipLaunchKernelGGL(gather_rows_kernel, dim3(srows), dim3(64), 0, s, xlp, (size_t)dim * elem, step, srows, slp,
                       d.fp8_scale, ssc, d.space == 0 ? xnm : nullptr, snm);
    LowpGemmArgs ga{};
    ga.A = qlp;
    ga.a_scale = qsc;
    ga.a_norm = qnm;
    ga.nq = nq;
    ga.dim = dim;
    ga.space = d.space;
    ga.B = slp;
    ga.b_scale = ssc;
    ga.b_norm = snm;
    ga.n = srows;
    ga.out_mat = smat;
    ga.ld_out = srows;
    ga.tiles_m = (nq + GM - 1) / GM;
    ga.tiles_n = (srows + GN - 1) / GN;
    launch_gemm(ga, precision, false, variant, s);
    // ~1.2 c candidates expected (buffer holds 4c); never below k + 1 sample ranks
    double want = 1.2 * (double)c * (double)srows / (double)n;
    uint32_t rank = (uint32_t)std::min<double>((double)srows, std::max<double>((double)k + 1, std::ceil(want)));
    hipLaunchKernelGGL(kth_kernel, dim3(nq), dim3(256), 0, s, smat, (size_t)srows, srows, rank, tau);
    hipLaunchKernelGGL(fill_u32_kernel, dim3((2 * nq + 255) / 256), dim3(256), 0, s, ccnt, 2 * nq, 0u);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, (uint32_t*)tie_ctr, 16u, 0u);
    (void)hipEventRecord(ev.ev[1], s);
    // filter pass over every row
    ga.B = xlp;
    ga.b_scale = d.fp8_scale;
    ga.b_norm = xnm;
    ga.n = n;
    ga.tau = tau;
    ga.cand_rows = crow;
    ga.cand_d = cd;
    ga.cand_cnt = ccnt;
    ga.cap = RR_CAP;
    launch_gemm(ga, precision, true, variant, s);
    (void)hipEventRecord(ev.ev[2], s);
    uint32_t cpow = 2;
    while (cpow < c) cpow <<= 1;
    size_t lds = (size_t)std::max<uint32_t>(cpow, RR_TIES) * 8 + (size_t)((dim + 1) & ~1u) * 4;
    const float* bnd = xstats ? bound : nullptr;
    unsigned long long* tc = tie_rel > 0 ? tie_ctr : nullptr;
    if (d.space == 0) {
        (void)hipFuncSetAttribute((const void*)rerank_kernel<0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(rerank_kernel<0>, dim3(nq), dim3(RR_THREADS), lds, s, q, d.vec, d.labels, dim, crow, cd,
                           ccnt, (uint32_t)RR_CAP, c, k, labels, dists, (size_t)k, ovf, n, (const float*)tau, bnd, tc,
                           tie_rel, tie_ctr + 4);
    } else {
        (void)hipFuncSetAttribute((const void*)rerank_kernel<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(rerank_kernel<1>, dim3(nq), dim3(RR_THREADS), lds, s, q, d.vec, d.labels, dim, crow, cd,
                           ccnt, (uint32_t)RR_CAP, c, k, labels, dists, (size_t)k, ovf, n, (const float*)tau, bnd, tc,
                           tie_rel, tie_ctr + 4);
    }
    (void)hipEventRecord(ev.ev[3], s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // queries the re-rank flagged take the exact path, all in one batch
    std::vector<uint32_t> flags(nq);
    unsigned long long ties[5] = {0, 0, 0, 0, 0};
    float bmax = 0.f;
    e = hipMemcpyAsync(flags.data(), ovf, (size_t)nq * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && rep) e = hipMemcpyAsync(ties, tie_ctr, sizeof(ties), hipMemcpyDeviceToHost, s);
    std::vector<float> bounds(xstats && rep ? nq : 0);
    if (e == hipSuccess && !bounds.empty()) e = hipMemcpyAsync(bounds.data(), bound, (size_t)nq * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    for (float v : bounds) bmax = std::max(bmax, v);
    std::vector<uint32_t> idx;
    uint32_t n_ovf = 0, n_unc = 0;
    for (uint32_t i = 0; i < nq; i++) {
        if (!flags[i]) continue;
        idx.push_back(i);
        if (flags[i] == 2)
            n_unc++;
        else
            n_ovf++;
    }
    if (!idx.empty()) {
        // the candidate region is free now: the list of queries, their vectors and the exact
        // rows (nf * (4 + 4 dim + 12 k) bytes <= nq * 8 * RR_CAP for dim <= 8192, k <= 2048);
        // the sample matrix region holds the distance chunk (at least one query row of n)
        const uint32_t nf = (uint32_t)idx.size();
        char* cp = (char*)crow;
        uint32_t* didx = (uint32_t*)cp;
        float* fq = (float*)(cp + align_up((size_t)nf * 4));
        uint64_t* fl = (uint64_t*)((char*)fq + align_up((size_t)nf * dim * 4));
        float* fd = (float*)((char*)fl + align_up((size_t)nf * k * 8));
        if ((char*)fd + (size_t)nf * k * 4 > cp + (size_t)nq * RR_CAP * 8) return hipErrorInvalidValue;
        e = hipMemcpyAsync(didx, idx.data(), (size_t)nf * 4, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(gather_queries_kernel, dim3(nf), dim3(256), 0, s, q, dim, (const uint32_t*)didx, nf, fq);
        e = launch_bf_exact(d, fq, nf, k, fl, fd, smat, s, smat_bytes);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(scatter_rows_kernel, dim3(nf), dim3(256), 0, s, (const uint64_t*)fl, (const float*)fd, k,
                           (const uint32_t*)didx, nf, labels, dists);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        // the host list must outlive the copy
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
    }
    if (stats) {
        float t0 = 0, t1 = 0, t2 = 0;
        (void)hipEventSynchronize(ev.ev[3]);
        (void)hipEventElapsedTime(&t0, ev.ev[0], ev.ev[1]);
        (void)hipEventElapsedTime(&t1, ev.ev[1], ev.ev[2]);
        (void)hipEventElapsedTime(&t2, ev.ev[2], ev.ev[3]);
        stats[0] = t1;  // filter GEMM over all rows
        stats[1] = t0;  // conversion + sample GEMM + thresholds
        stats[2] = t2;  // re-rank
        stats[3] = (double)idx.size();
    }
    if (rep) {
        rep->exact_overflow = n_ovf;
        rep->exact_uncertified = n_unc;
        rep->near_ties = ties[0];
        rep->near_tie_disagree = ties[1];
        rep->boundary_near_ties = ties[2];
        rep->boundary_disagree = ties[3];
        rep->max_bound = bmax;
        rep->certified = xstats != nullptr;
        rep->reranked_rows = ties[4];
    }
    return hipGetLastError();
}

// FP64 check of the final lists (see hg_bf_fp64_check).  The row of each listed label is
// found by binary search in (sorted_labels, sorted_rows), the index's labels in ascending order;
// a label the index does not hold counts in counters[2] and is skipped.
__device__ __forceinline__ int64_t row_of_label(const uint64_t* sl, const uint32_t* sr, uint32_t n, uint64_t label) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sl[mid] < label)
            lo = mid + 1;
        else
            hi = mid;
    }
    return (lo < n && sl[lo] == label) ? (int64_t)sr[lo] : -1;
}

__global__ void fp64_check_kernel(int space, const float* __restrict__ Q, const float* __restrict__ X,
                                  const uint64_t* __restrict__ sorted_labels, const uint32_t* __restrict__ sorted_rows,
                                  uint32_t n, uint32_t dim, uint32_t nq, uint32_t k, const uint64_t* __restrict__ labels,
                                  const float* __restrict__ dists, double rel_tol, unsigned long long* counters,
                                  double* maxes) {
    uint32_t q = blockIdx.x;
    if (q >= nq) return;
    __shared__ double d64[2048];
    for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
        const uint64_t li = labels[(size_t)q * k + i];
        const int64_t r = li == ~0ull ? -1 : row_of_label(sorted_labels, sorted_rows, n, li);
        if (r < 0) {
            if (li != ~0ull) atomicAdd(&counters[2], 1ull);  // an empty slot is no error
            d64[i] = NAN;
            continue;
        }
        double s = 0;
        const float* a = Q + (size_t)q * dim;
        const float* x = X + (size_t)r * dim;
        if (space == 0) {
            for (uint32_t j = 0; j < dim; j++) {
                double t = (double)a[j] - (double)x[j];
                s += t * t;
            }
        } else {
            for (uint32_t j = 0; j < dim; j++) s += (double)a[j] * (double)x[j];
            s = 1.0 - s;
        }
        d64[i] = s;
        double err = fabs((double)dists[(size_t)q * k + i] - s);
        double rel = err / fmax(fabs(s), 1e-30);
        unsigned long long eb = (unsigned long long)__double_as_longlong(err);
        unsigned long long rb = (unsigned long long)__double_as_longlong(rel);
        atomicMax((unsigned long long*)&maxes[0], eb);
        atomicMax((unsigned long long*)&maxes[1], rb);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i + 1 < k; i += blockDim.x) {
        float a = dists[(size_t)q * k + i], b = dists[(size_t)q * k + i + 1];
        if ((double)b - (double)a <= rel_tol * fabs((double)a)) {
            atomicAdd(&counters[0], 1ull);
            // the lists order equal FP32 distances by label; FP64 disagrees when it puts the
            // later entry strictly first, or ties them against that label order
            const uint64_t la = labels[(size_t)q * k + i], lb = labels[(size_t)q * k + i + 1];
            if (d64[i] > d64[i + 1] || (d64[i] == d64[i + 1] && la > lb)) atomicAdd(&counters[1], 1ull);
        }
    }
}

hipError_t launch_fp64_check(const BfData& d, const float* q, uint32_t nq, uint32_t k, const uint64_t* labels,
                             const float* dists, double rel_tol, const uint64_t* sorted_labels,
                             const uint32_t* sorted_rows, double* out_host, void* scratch, hipStream_t s) {
    if (k > 2048) return hipErrorInvalidValue;
    unsigned long long* ctr = (unsigned long long*)scratch;  // [0] pairs, [1] disagreements, [2] unknown labels
    double* mx = (double*)((char*)scratch + 32);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, (uint32_t*)scratch, 12u, 0u);
    hipLaunchKernelGGL(fp64_check_kernel, dim3(nq), dim3(256), 0, s, (int)d.space, q, d.vec, sorted_labels,
                       sorted_rows, d.n, d.dim, nq, k, labels, dists, rel_tol, ctr, mx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    unsigned long long hc[3];
    double hm[2];
    e = hipMemcpyAsync(hc, ctr, 24, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(hm, mx, 16, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    out_host[0] = (double)hc[0];
    out_host[1] = (double)hc[1];
    out_host[2] = hm[0];
    out_host[3] = hm[1];
    out_host[4] = (double)hc[2];
    return hipSuccess;
}

}  // namespace hg
