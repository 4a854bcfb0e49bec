// Device distance functions in the reference's exact FP32 order (shared by the kernels).
#pragma once
#include <hip/hip_runtime.h>

namespace hg {

__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Generic dimensions (dim % 16 != 0): one lane per vector, the reference kernel for that
// dimension class written out in scalar form (space_l2.h:6-20,149-205; space_ip.h:7-19,24-69,
// 307-332; selection space_l2.h:227-235, space_ip.h:372-380).
// Scalar L2 loop as the reference's compiler emits it: rounded squares of groups of four added
// in order, the n % 4 remainder fused.
__device__ __forceinline__ float l2_scalar(const float* q, const float* x, int n) {
    float r = 0.f;
    int nv = n / 4 * 4;
    for (int i = 0; i < nv; i++) {
        float t = q[i] - x[i];
        float sq = t * t;
        r = r + sq;
    }
    for (int i = nv; i < n; i++) {
        float t = q[i] - x[i];
        r = fma_(t, t, r);
    }
    return r;
}
__device__ __forceinline__ float ip_scalar(const float* q, const float* x, int n) {
    float r = 0.f;
    for (int i = 0; i < n; i++) r = fma_(q[i], x[i], r);
    return r;
}
__device__ __forceinline__ float l2_lanes16(const float* q, const float* x, int n16) {
    float acc[16];
#pragma unroll
    for (int l = 0; l < 16; l++) acc[l] = 0.f;
    for (int c = 0; c < n16; c += 16)
#pragma unroll
        for (int l = 0; l < 16; l++) {
            float d = q[c + l] - x[c + l];
            acc[l] = fma_(d, d, acc[l]);
        }
    float r = acc[0];
#pragma unroll
    for (int l = 1; l < 16; l++) r = r + acc[l];
    return r;
}
__device__ __forceinline__ float l2_lanes4(const float* q, const float* x, int n4) {
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int c = 0; c < n4; c += 4)
#pragma unroll
        for (int l = 0; l < 4; l++) {
            float d = q[c + l] - x[c + l];
            acc[l] = fma_(d, d, acc[l]);
        }
    return ((acc[0] + acc[1]) + acc[2]) + acc[3];
}
__device__ __forceinline__ float ip_lanes16(const float* q, const float* x, int n16) {
    float acc[16];
#pragma unroll
    for (int l = 0; l < 16; l++) acc[l] = 0.f;
    for (int c = 0; c < n16; c += 16)
#pragma unroll
        for (int l = 0; l < 16; l++) acc[l] = fma_(q[c + l], x[c + l], acc[l]);
    float h8[8], h4[4];
#pragma unroll
    for (int i = 0; i < 8; i++) h8[i] = acc[i] + acc[i + 8];
#pragma unroll
    for (int i = 0; i < 4; i++) h4[i] = h8[i] + h8[i + 4];
    float a = h4[0] + h4[2], b = h4[1] + h4[3];
    return a + b;
}
__device__ __forceinline__ float ip_lanes8_4(const float* q, const float* x, int n) {
    int n16 = n / 16 * 16, n4 = n / 4 * 4;
    float acc8[8];
#pragma unroll
    for (int l = 0; l < 8; l++) acc8[l] = 0.f;
    for (int c = 0; c < n16; c += 8)
#pragma unroll
        for (int l = 0; l < 8; l++) acc8[l] = fma_(q[c + l], x[c + l], acc8[l]);
    float acc[4];
#pragma unroll
    for (int l = 0; l < 4; l++) acc[l] = acc8[l] + acc8[l + 4];
    for (int c = n16; c < n4; c += 4)
#pragma unroll
        for (int l = 0; l < 4; l++) acc[l] = fma_(q[c + l], x[c + l], acc[l]);
    return ((acc[0] + acc[1]) + acc[2]) + acc[3];
}
__device__ __forceinline__ float dist_generic(int space, const float* q, const float* x, int dim) {
    if (space == 0) {
        if (dim % 16 == 0) return l2_lanes16(q, x, dim);
        if (dim % 4 == 0) return l2_lanes4(q, x, dim);
        if (dim > 16) {
            int n16 = dim >> 4 << 4;
            float r = l2_lanes16(q, x, n16);
            float t = l2_scalar(q + n16, x + n16, dim - n16);
            return r + t;
        }
        if (dim > 4) {
            int n4 = dim >> 2 << 2;
            float r = l2_lanes4(q, x, n4);
            float t = l2_scalar(q + n4, x + n4, dim - n4);
            return r + t;
        }
        return l2_scalar(q, x, dim);
    }
    if (dim % 16 == 0) return 1.0f - ip_lanes16(q, x, dim);
    if (dim % 4 == 0) return 1.0f - ip_lanes8_4(q, x, dim);
    if (dim > 16) {
        int n16 = dim >> 4 << 4;
        float r = ip_lanes16(q, x, n16);
        float t = ip_scalar(q + n16, x + n16, dim - n16);
        return 1.0f - (r + t);
    }
    if (dim > 4) {
        int n4 = dim >> 2 << 2;
        float r = ip_lanes8_4(q, x, n4);
        float t = ip_scalar(q + n4, x + n4, dim - n4);
        return 1.0f - (r + t);
    }
    return 1.0f - ip_scalar(q, x, dim);
}

// ------------------------------------------------------------------------------------------
// Distances.  dim % 16 == 0: four lanes per vector, lane g of a group owns the reference's
// AVX-512 accumulator lanes 4g..4g+3 and reads coordinates 16t+4g..16t+4g+3 with one 16-byte
// load per chunk t (coalesced 64 bytes per group and chunk).
//   L2 (L2SqrSIMD16ExtAVX512, space_l2.h:25-54; GCC fuses the mul+add): acc = fma(d, d, acc),
//      then the 16 accumulators summed left to right -- carried across the four lanes.
//   IP (InnerProductSIMD16ExtAVX512, space_ip.h:146-204): acc = fma(q, x, acc), then the
//      _mm512_reduce_add_ps halving tree (i,i+8), (i,i+4), (i,i+2), (0,1); distance 1 - sum.
// Results: out[idx] for each of the c ids in ids[].
// ------------------------------------------------------------------------------------------
// One chunk of four coordinates into the lane's four accumulators.
template <int SPACE>
__device__ __forceinline__ void group16_acc(const float4& x, const float4& q, float& a0, float& a1, float& a2,
                                            float& a3) {
    if (SPACE == 0) {
        float d0 = q.x - x.x, d1 = q.y - x.y, d2 = q.z - x.z, d3 = q.w - x.w;
        a0 = fma_(d0, d0, a0);
        a1 = fma_(d1, d1, a1);
        a2 = fma_(d2, d2, a2);
        a3 = fma_(d3, d3, a3);
    } else {
        a0 = fma_(q.x, x.x, a0);
        a1 = fma_(q.y, x.y, a1);
        a2 = fma_(q.z, x.z, a2);
        a3 = fma_(q.w, x.w, a3);
    }
}

// The group's 16 accumulators (four per lane) to the distance, in the reference's order; every
// lane of the group returns it.
template <int SPACE>
__device__ __forceinline__ float group16_reduce(float a0, float a1, float a2, float a3, int g, int lane) {
    const int gbase = lane & ~3;
    if (SPACE == 0) {
        float r = ((a0 + a1) + a2) + a3;
#pragma unroll
        for (int s = 1; s < 4; s++) {
            float p = __shfl(r, gbase + s - 1);
            if (g == s) r = (((p + a0) + a1) + a2) + a3;
        }
        return __shfl(r, gbase + 3);
    }
    // (i, i+8): lanes g=0,1 pair with g=2,3
    float o0 = __shfl(a0, lane + 2), o1 = __shfl(a1, lane + 2), o2 = __shfl(a2, lane + 2), o3 = __shfl(a3, lane + 2);
    float h0 = a0 + o0, h1 = a1 + o1, h2 = a2 + o2, h3 = a3 + o3;
    // (i, i+4): lane g=0 pairs with g=1
    float p0 = __shfl(h0, lane + 1), p1 = __shfl(h1, lane + 1), p2 = __shfl(h2, lane + 1), p3 = __shfl(h3, lane + 1);
    float u0 = h0 + p0, u1 = h1 + p1, u2 = h2 + p2, u3 = h3 + p3;
    float s0 = u0 + u2, s1 = u1 + u3;
    float ip = s0 + s1;
    return __shfl(1.0f - ip, gbase);
}

template <int SPACE, int NCH>
__device__ __forceinline__ void dist_group16(const float* __restrict__ row, const float* qs, int nch, bool act, int g,
                                             float& res, int lane) {
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    const float4* r4 = reinterpret_cast<const float4*>(row);
    const int n = NCH > 0 ? NCH : nch;
#pragma unroll
    for (int t = 0; t < (NCH > 0 ? NCH : 1); t++) {
        if (NCH == 0) break;
        float4 x = act ? r4[4 * t + g] : make_float4(0.f, 0.f, 0.f, 0.f);
        group16_acc<SPACE>(x, *reinterpret_cast<const float4*>(qs + 16 * t + 4 * g), a0, a1, a2, a3);
    }
    if (NCH == 0) {
#pragma unroll 4
        for (int t = 0; t < n; t++) {
            float4 x = act ? r4[4 * t + g] : make_float4(0.f, 0.f, 0.f, 0.f);
            group16_acc<SPACE>(x, *reinterpret_cast<const float4*>(qs + 16 * t + 4 * g), a0, a1, a2, a3);
        }
    }
    res = group16_reduce<SPACE>(a0, a1, a2, a3, g, lane);
}

// The same with the row's NCH chunks already in registers (x[t] = coordinates 16t+4g..16t+4g+3).
template <int SPACE, int NCH>
__device__ __forceinline__ float dist_group16_regs(const float4 (&x)[NCH], const float* qs, int g, int lane) {
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
#pragma unroll
    for (int t = 0; t < NCH; t++)
        group16_acc<SPACE>(x[t], *reinterpret_cast<const float4*>(qs + 16 * t + 4 * g), a0, a1, a2, a3);
    return group16_reduce<SPACE>(a0, a1, a2, a3, g, lane);
}

// The same from the 8-bit coded row: w[t] holds coordinates 16t+4g..16t+4g+3 as four bytes (low
// byte first).  Every coded value is an integer in 0..255 whose FP32 form is the stored
// coordinate itself, so the conversion is exact and the accumulation sees the FP32 values.
template <int SPACE, int NCH>
__device__ __forceinline__ float dist_group16_u8(const uint32_t (&w)[NCH], const float* qs, int g, int lane) {
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
#pragma unroll
    for (int t = 0; t < NCH; t++) {
        const uint32_t v = w[t];
        const float4 x = make_float4((float)(v & 0xFFu), (float)((v >> 8) & 0xFFu), (float)((v >> 16) & 0xFFu),
                                     (float)(v >> 24));
        group16_acc<SPACE>(x, *reinterpret_cast<const float4*>(qs + 16 * t + 4 * g), a0, a1, a2, a3);
    }
    return group16_reduce<SPACE>(a0, a1, a2, a3, g, lane);
}

}  // namespace hg
