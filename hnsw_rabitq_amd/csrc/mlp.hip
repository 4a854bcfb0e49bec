// Learned early-termination network: forward, backward and Adam on gfx950.
//
// Network: in -> hidden -> hidden -> 1 with tanh-form GELU, weights used in BF16 (rounded from
// an FP32 master copy), FP32 accumulation -- the same arithmetic as the in-search inference
// (search.hip, mlp_logit).  Training minimises mean binary cross-entropy with logits; the
// gradient flows to the FP32 masters through the rounding (straight-through).
//
// Layout of the parameter vector: W1[h][in], b1[h], W2[h][h], b2[h], w3[h], b3.
#include "common.h"
#include "dist.h"
#include "kernels.h"

namespace hg {

constexpr int MLP_MAX_IN = 32, MLP_MAX_H = 64;

__device__ __forceinline__ float bf16r(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xFFFF0000u);
}
__device__ __forceinline__ float gelu_f(float x) {
    const float c = 0.7978845608028654f;
    return 0.5f * x * (1.0f + tanhf(c * (x + 0.044715f * x * x * x)));
}
__device__ __forceinline__ float gelu_grad_f(float x) {
    const float c = 0.7978845608028654f;
    float u = c * (x + 0.044715f * x * x * x);
    float t = tanhf(u);
    float du = c * (1.0f + 3.0f * 0.044715f * x * x);
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * du;
}

struct MlpShape {
    int in, h;
    __device__ __forceinline__ int oW1() const { return 0; }
    __device__ __forceinline__ int ob1() const { return h * in; }
    __device__ __forceinline__ int oW2() const { return h * in + h; }
    __device__ __forceinline__ int ob2() const { return h * in + h + h * h; }
    __device__ __forceinline__ int ow3() const { return h * in + 2 * h + h * h; }
    __device__ __forceinline__ int ob3() const { return h * in + 3 * h + h * h; }
    __device__ __forceinline__ int count() const { return h * in + 3 * h + h * h + 1; }
};

// Effective (BF16-rounded) weights, FP32 biases, staged in LDS.
__device__ void stage_params(const float* __restrict__ P, MlpShape s, float* w) {
    const int n = s.count();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        bool is_bias = (i >= s.ob1() && i < s.oW2()) || (i >= s.ob2() && i < s.ow3()) || i == s.ob3();
        w[i] = is_bias ? P[i] : bf16r(P[i]);
    }
    __syncthreads();
}

// One thread per sample.  Writes logits; when y != null also writes the per-sample backward
// quantities (d1 [n][h], d2 [n][h], h1 [n][h], h2 [n][h], dout [n]) and per-block loss sums.
__global__ __launch_bounds__(256) void mlp_fwd_kernel(const float* __restrict__ P, int in, int hdim,
                                                      const float* __restrict__ X, const float* __restrict__ Y,
                                                      uint32_t n, float* __restrict__ logits,
                                                      float* __restrict__ H1, float* __restrict__ H2,
                                                      float* __restrict__ D1, float* __restrict__ D2,
                                                      float* __restrict__ DO, float* __restrict__ block_loss) {
    extern __shared__ float w[];
    __shared__ float red[256];
    MlpShape s{in, hdim};
    stage_params(P, s, w);
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    float loss = 0.f;
    if (r < n) {
        float x[MLP_MAX_IN], a1[MLP_MAX_H], h1[MLP_MAX_H], a2[MLP_MAX_H];
        for (int i = 0; i < in; i++) x[i] = X[(size_t)r * in + i];
        for (int j = 0; j < hdim; j++) {
            float acc = w[s.ob1() + j];
            for (int i = 0; i < in; i++) acc = fma_(w[s.oW1() + j * in + i], x[i], acc);
            a1[j] = acc;
            h1[j] = gelu_f(acc);
        }
        float o = w[s.ob3()];
        for (int j = 0; j < hdim; j++) {
            float acc = w[s.ob2() + j];
            for (int i = 0; i < hdim; i++) acc = fma_(w[s.oW2() + j * hdim + i], h1[i], acc);
            a2[j] = acc;
        }
        // output: bias + sum_j w3[j] * gelu(a2[j]) accumulated in j order
        float acc_o = 0.f;
        for (int j = 0; j < hdim; j++) acc_o = fma_(gelu_f(a2[j]), w[s.ow3() + j], acc_o);
        o = acc_o + o;
        logits[r] = o;
        if (Y) {
            float y = Y[r];
            float ao = fabsf(o);
            loss = fmaxf(o, 0.f) - o * y + log1pf(expf(-ao));
            float sig = 1.0f / (1.0f + expf(-o));
            float dout = (sig - y) / (float)n;
            DO[r] = dout;
            float d2[MLP_MAX_H];
            for (int j = 0; j < hdim; j++) {
                float h2v = gelu_f(a2[j]);
                H2[(size_t)r * hdim + j] = h2v;
                d2[j] = dout * w[s.ow3() + j] * gelu_grad_f(a2[j]);
                D2[(size_t)r * hdim + j] = d2[j];
                H1[(size_t)r * hdim + j] = h1[j];
            }
            for (int i = 0; i < hdim; i++) {
                float acc = 0.f;
                for (int j = 0; j < hdim; j++) acc = fma_(d2[j], w[s.oW2() + j * hdim + i], acc);
                D1[(size_t)r * hdim + i] = acc * gelu_grad_f(a1[i]);
            }
        }
    }
    if (Y) {
        red[threadIdx.x] = loss;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) block_loss[blockIdx.x] = red[0];
    }
}

// One thread per parameter: gradient as an ordered sum over the minibatch, then Adam.
__global__ __launch_bounds__(256) void mlp_grad_adam_kernel(float* __restrict__ P, float* __restrict__ M,
                                                            float* __restrict__ V, int step, int in, int hdim,
                                                            const float* __restrict__ X, uint32_t n,
                                                            const float* __restrict__ H1, const float* __restrict__ H2,
                                                            const float* __restrict__ D1, const float* __restrict__ D2,
                                                            const float* __restrict__ DO, float lr, float b1, float b2,
                                                            float eps, float* __restrict__ grad_out) {
    MlpShape s{in, hdim};
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= s.count()) return;
    float g = 0.f;
    if (p < s.ob1()) {
        int j = p / in, i = p % in;
        for (uint32_t r = 0; r < n; r++) g += D1[(size_t)r * hdim + j] * X[(size_t)r * in + i];
    } else if (p < s.oW2()) {
        int j = p - s.ob1();
        for (uint32_t r = 0; r < n; r++) g += D1[(size_t)r * hdim + j];
    } else if (p < s.ob2()) {
        int q = p - s.oW2();
        int j = q / hdim, i = q % hdim;
        for (uint32_t r = 0; r < n; r++) g += D2[(size_t)r * hdim + j] * H1[(size_t)r * hdim + i];
    } else if (p < s.ow3()) {
        int j = p - s.ob2();
        for (uint32_t r = 0; r < n; r++) g += D2[(size_t)r * hdim + j];
    } else if (p < s.ob3()) {
        int j = p - s.ow3();
        for (uint32_t r = 0; r < n; r++) g += DO[r] * H2[(size_t)r * hdim + j];
    } else {
        for (uint32_t r = 0; r < n; r++) g += DO[r];
    }
    if (grad_out) grad_out[p] = g;
    if (lr != 0.f) {
        float m = b1 * M[p] + (1.f - b1) * g;
        float v = b2 * V[p] + (1.f - b2) * g * g;
        M[p] = m;
        V[p] = v;
        float mh = m / (1.f - powf(b1, (float)step));
        float vh = v / (1.f - powf(b2, (float)step));
        P[p] -= lr * mh / (sqrtf(vh) + eps);
    }
}

__global__ void sum_kernel(const float* __restrict__ x, uint32_t n, float* __restrict__ out) {
    __shared__ float red[256];
    float s = 0.f;
    for (uint32_t i = threadIdx.x; i < n; i += 256) s += x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

static size_t mlp_count(int in, int h) { return (size_t)h * in + 3 * (size_t)h + (size_t)h * h + 1; }

hipError_t launch_mlp_forward(const float* params, int in_dim, int hidden, const float* x, uint32_t n, float* logits,
                              hipStream_t s) {
    if (in_dim > MLP_MAX_IN || hidden > MLP_MAX_H) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    size_t lds = mlp_count(in_dim, hidden) * 4;
    hipLaunchKernelGGL(mlp_fwd_kernel, dim3((n + 255) / 256), dim3(256), lds, s, params, in_dim, hidden, x,
                       (const float*)nullptr, n, logits, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    return hipGetLastError();
}

// Scratch layout (device, provided by caller in `grad_dev`'s companion buffer): see capi_stop.
hipError_t launch_mlp_train_step(float* params, float* m, float* v, int step, int in_dim, int hidden, const float* x,
                                 const float* y, uint32_t n, float lr, float b1, float b2, float eps, float* loss_dev,
                                 float* grad_dev, hipStream_t s) {
    if (in_dim > MLP_MAX_IN || hidden > MLP_MAX_H) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    // loss_dev points at a scratch area: [1 loss][blocks][n logits][n dout][4 * n * hidden]
    const uint32_t blocks = (n + 255) / 256;
    float* loss = loss_dev;
    float* bl = loss_dev + 1;
    float* lg = bl + blocks;
    float* dout = lg + n;
    float* H1 = dout + n;
    float* H2 = H1 + (size_t)n * hidden;
    float* D1 = H2 + (size_t)n * hidden;
    float* D2 = D1 + (size_t)n * hidden;
    size_t lds = mlp_count(in_dim, hidden) * 4;
    hipLaunchKernelGGL(mlp_fwd_kernel, dim3(blocks), dim3(256), lds, s, params, in_dim, hidden, x, y, n, lg, H1, H2,
                       D1, D2, dout, bl);
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, bl, blocks, loss);
    const int P = (int)mlp_count(in_dim, hidden);
    hipLaunchKernelGGL(mlp_grad_adam_kernel, dim3((P + 255) / 256), dim3(256), 0, s, params, m, v, step, in_dim,
                       hidden, x, n, H1, H2, D1, D2, dout, lr, b1, b2, eps, grad_dev);
    return hipGetLastError();
}

}  // namespace hg
