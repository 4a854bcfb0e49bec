// Host distance functions with the reference's exact floating-point order.
//
// The reference picks a kernel per dimension at space construction (space_l2.h:214-238,
// space_ip.h:348-383).  On an AVX-512 host with GCC -O3 -march=native (the Python extension
// flags, setup.py:76-103) the intrinsic mul+add pairs are contracted into FMA, so the bits are:
//   L2, dim%16==0 : 16 lanes acc[l] = fma(d,d,acc[l]); lanes summed left to right
//                   (L2SqrSIMD16ExtAVX512, space_l2.h:25-54)
//   L2, dim%4==0  : the same with 4 lanes (L2SqrSIMD4Ext, space_l2.h:165-186)
//   IP, dim%16==0 : 16 lanes of fma, GCC's _mm512_reduce_add_ps halving tree, 1 - sum
//                   (InnerProductSIMD16ExtAVX512, space_ip.h:146-204)
//   IP, dim%4==0  : 8 lanes over chunks of 16, high half folded, 4 lanes over chunks of 4,
//                   lanes summed left to right, 1 - sum (InnerProductSIMD4ExtAVX, space_ip.h:24-69)
//   residual forms: SIMD part + scalar tail (space_l2.h:149-205, space_ip.h:307-332); the L2
//                   tail is vectorised four wide by GCC (rounded squares, in-order adds) with a
//                   fused remainder.
// This file is compiled with -ffp-contract=off, so only the explicit fmaf calls fuse.
#pragma once
#include <cmath>
#include <cstddef>

namespace hnswgpu {

enum Space { SPACE_L2 = 0, SPACE_IP = 1 };

namespace dist_detail {

// Scalar L2 loop as compiled by the reference's flags: squares of groups of four rounded
// separately and added in order, the n % 4 remainder fused.
inline float l2_scalar(const float* a, const float* b, size_t n) {
    float res = 0.f;
    size_t nv = n / 4 * 4;
    for (size_t i = 0; i < nv; i++) {
        float t = a[i] - b[i];
        float sq = t * t;
        res = res + sq;
    }
    for (size_t i = nv; i < n; i++) {
        float t = a[i] - b[i];
        res = fmaf(t, t, res);
    }
    return res;
}

inline float ip_scalar(const float* a, const float* b, size_t n) {
    float res = 0.f;
    for (size_t i = 0; i < n; i++) res = fmaf(a[i], b[i], res);
    return res;
}

inline float l2_lanes16(const float* __restrict a, const float* __restrict b, size_t n16) {
    float acc[16];
    for (int l = 0; l < 16; l++) acc[l] = 0.f;
    for (size_t c = 0; c < n16; c += 16) {
#pragma GCC ivdep
        for (int l = 0; l < 16; l++) {
            float d = a[c + l] - b[c + l];
            acc[l] = fmaf(d, d, acc[l]);
        }
    }
    float r = acc[0];
    for (int l = 1; l < 16; l++) r += acc[l];
    return r;
}

inline float l2_lanes4(const float* a, const float* b, size_t n4) {
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (size_t c = 0; c < n4; c += 4)
        for (int l = 0; l < 4; l++) {
            float d = a[c + l] - b[c + l];
            acc[l] = fmaf(d, d, acc[l]);
        }
    return ((acc[0] + acc[1]) + acc[2]) + acc[3];
}

inline float ip_lanes16(const float* __restrict a, const float* __restrict b, size_t n16) {
    float acc[16];
    for (int l = 0; l < 16; l++) acc[l] = 0.f;
    for (size_t c = 0; c < n16; c += 16) {
#pragma GCC ivdep
        for (int l = 0; l < 16; l++) acc[l] = fmaf(a[c + l], b[c + l], acc[l]);
    }
    float h8[8], h4[4];
    for (int i = 0; i < 8; i++) h8[i] = acc[i] + acc[i + 8];
    for (int i = 0; i < 4; i++) h4[i] = h8[i] + h8[i + 4];
    float h20 = h4[0] + h4[2], h21 = h4[1] + h4[3];
    return h20 + h21;
}

inline float ip_lanes8_4(const float* a, const float* b, size_t n) {
    size_t n16 = n / 16 * 16, n4 = n / 4 * 4;
    float acc8[8];
    for (int l = 0; l < 8; l++) acc8[l] = 0.f;
    for (size_t c = 0; c < n16; c += 8)
        for (int l = 0; l < 8; l++) acc8[l] = fmaf(a[c + l], b[c + l], acc8[l]);
    float acc[4];
    for (int l = 0; l < 4; l++) acc[l] = acc8[l] + acc8[l + 4];
    for (size_t c = n16; c < n4; c += 4)
        for (int l = 0; l < 4; l++) acc[l] = fmaf(a[c + l], b[c + l], acc[l]);
    return ((acc[0] + acc[1]) + acc[2]) + acc[3];
}

}  // namespace dist_detail

inline float distance_l2(const float* a, const float* b, size_t dim) {
    using namespace dist_detail;
    if (dim % 16 == 0) return l2_lanes16(a, b, dim);
    if (dim % 4 == 0) return l2_lanes4(a, b, dim);
    if (dim > 16) {
        size_t n16 = dim >> 4 << 4;
        float res = l2_lanes16(a, b, n16);
        float tail = l2_scalar(a + n16, b + n16, dim - n16);
        return res + tail;
    }
    if (dim > 4) {
        size_t n4 = dim >> 2 << 2;
        float res = l2_lanes4(a, b, n4);
        float tail = l2_scalar(a + n4, b + n4, dim - n4);
        return res + tail;
    }
    return l2_scalar(a, b, dim);
}

inline float distance_ip(const float* a, const float* b, size_t dim) {
    using namespace dist_detail;
    if (dim % 16 == 0) return 1.0f - ip_lanes16(a, b, dim);
    if (dim % 4 == 0) return 1.0f - ip_lanes8_4(a, b, dim);
    if (dim > 16) {
        size_t n16 = dim >> 4 << 4;
        float res = ip_lanes16(a, b, n16);
        float tail = ip_scalar(a + n16, b + n16, dim - n16);
        return 1.0f - (res + tail);
    }
    if (dim > 4) {
        size_t n4 = dim >> 2 << 2;
        float res = ip_lanes8_4(a, b, n4);
        float tail = ip_scalar(a + n4, b + n4, dim - n4);
        return 1.0f - (res + tail);
    }
    return 1.0f - ip_scalar(a, b, dim);
}

typedef float (*DistFn)(const float*, const float*, size_t);

inline DistFn distance_fn(int space) { return space == SPACE_L2 ? distance_l2 : distance_ip; }

// Index.normalize_vector (bindings.cpp:241-248): the cosine space stores and queries
// unit-normalised vectors in the inner-product space.  The squared norm follows the same
// compiled form as the scalar L2 loop: rounded squares of groups of four added in order, the
// dim % 4 remainder fused.
inline void normalize_vector(const float* in, float* out, size_t dim) {
    float norm = 0.0f;
    size_t nv = dim / 4 * 4;
    for (size_t i = 0; i < nv; i++) {
        float sq = in[i] * in[i];
        norm = norm + sq;
    }
    for (size_t i = nv; i < dim; i++) norm = fmaf(in[i], in[i], norm);
    norm = 1.0f / (sqrtf(norm) + 1e-30f);
    for (size_t i = 0; i < dim; i++) out[i] = in[i] * norm;
}

}  // namespace hnswgpu
