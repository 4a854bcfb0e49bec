// C ABI: learned early-termination model (device parameters, training, inference settings).
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <vector>

#include "capi_common.h"
#include "kernels.h"

using namespace hgapi;

HG_API int hg_stop_model_create(int in_dim, int hidden, int device, void** out) {
    return guarded([&] {
        require(out != nullptr, "null output handle");
        require(in_dim == hg::kStopFeatures, "the search kernel produces exactly 12 traversal features");
        require(hidden >= 1 && hidden <= 64, "hidden width must be in [1, 64]");
        DeviceGuard dg(device);
        std::unique_ptr<StopModel> m(new StopModel());
        m->device = device;
        m->in_dim = in_dim;
        m->hidden = hidden;
        m->nparam = mlp_param_count(in_dim, hidden);
        m->params.reserve(m->nparam * 4);
        m->m.reserve(m->nparam * 4);
        m->v.reserve(m->nparam * 4);
        m->mean.reserve(in_dim * 4);
        m->inv_std.reserve(in_dim * 4);
        hip_check(hipMemset(m->params.p, 0, m->nparam * 4), "hipMemset");
        hip_check(hipMemset(m->m.p, 0, m->nparam * 4), "hipMemset");
        hip_check(hipMemset(m->v.p, 0, m->nparam * 4), "hipMemset");
        std::vector<float> zero(in_dim, 0.f), one(in_dim, 1.f);
        hip_check(hipMemcpy(m->mean.p, zero.data(), in_dim * 4, hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemcpy(m->inv_std.p, one.data(), in_dim * 4, hipMemcpyHostToDevice), "hipMemcpy");
        *out = m.release();
        return HG_OK;
    });
}

HG_API int hg_stop_model_free(void* mv) {
    delete static_cast<StopModel*>(mv);
    return HG_OK;
}

HG_API int hg_stop_model_param_count(void* mv, size_t* count) {
    return guarded([&] {
        require(mv && count, "null argument");
        *count = static_cast<StopModel*>(mv)->nparam;
        return HG_OK;
    });
}

HG_API int hg_stop_model_set_params(void* mv, const float* params, size_t count) {
    return guarded([&] {
        require(mv && params, "null argument");
        auto* m = static_cast<StopModel*>(mv);
        require(count == m->nparam, "parameter count mismatch");
        DeviceGuard dg(m->device);
        hip_check(hipMemcpy(m->params.p, params, count * 4, hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemset(m->m.p, 0, count * 4), "hipMemset");
        hip_check(hipMemset(m->v.p, 0, count * 4), "hipMemset");
        m->step = 0;
        return HG_OK;
    });
}

HG_API int hg_stop_model_get_params(void* mv, float* params, size_t count) {
    return guarded([&] {
        require(mv && params, "null argument");
        auto* m = static_cast<StopModel*>(mv);
        require(count == m->nparam, "parameter count mismatch");
        DeviceGuard dg(m->device);
        hip_check(hipMemcpy(params, m->params.p, count * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        return HG_OK;
    });
}

HG_API int hg_stop_model_set_norm(void* mv, const float* mean, const float* inv_std, size_t count) {
    return guarded([&] {
        require(mv && mean && inv_std, "null argument");
        auto* m = static_cast<StopModel*>(mv);
        require(count == (size_t)m->in_dim, "feature count mismatch");
        DeviceGuard dg(m->device);
        hip_check(hipMemcpy(m->mean.p, mean, count * 4, hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemcpy(m->inv_std.p, inv_std, count * 4, hipMemcpyHostToDevice), "hipMemcpy");
        return HG_OK;
    });
}

HG_API int hg_stop_model_config(void* mv, int period, float threshold, int min_hops) {
    return guarded([&] {
        require(mv != nullptr, "null model");
        require(period >= 1, "period must be >= 1");
        require(threshold > 0.f && threshold < 1.f, "threshold must be in (0, 1)");
        auto* m = static_cast<StopModel*>(mv);
        m->period = period;
        m->threshold = threshold;
        m->min_hops = min_hops < 0 ? 0 : min_hops;
        return HG_OK;
    });
}

// Extension of the learned search: where a query's search ends by its own rule at the given ef,
// the network is asked whether the top-k has converged; below ext_threshold the query is answered
// again by the search at ext_ef (ext_ef = 0 turns the extension off).
HG_API int hg_stop_model_config_extend(void* mv, int ext_ef, float ext_threshold) {
    return guarded([&] {
        require(mv != nullptr, "null model");
        require(ext_ef >= 0 && ext_ef <= 4096, "ext_ef must be in [0, 4096]");
        require(ext_threshold > 0.f && ext_threshold < 1.f, "ext_threshold must be in (0, 1)");
        auto* m = static_cast<StopModel*>(mv);
        m->ext_ef = ext_ef;
        m->ext_threshold = ext_threshold;
        return HG_OK;
    });
}

HG_API int hg_stop_model_forward(void* mv, const float* dx, size_t n, float* dlogits, void* stream) {
    return guarded([&] {
        require(mv && (n == 0 || (dx && dlogits)), "null argument");
        auto* m = static_cast<StopModel*>(mv);
        DeviceGuard dg(m->device);
        hip_check(hg::launch_mlp_forward((const float*)m->params.p, m->in_dim, m->hidden, dx, (uint32_t)n, dlogits,
                                         (hipStream_t)stream),
                  "mlp forward");
        return HG_OK;
    });
}

HG_API int hg_stop_model_train_step(void* mv, const float* dx, const float* dy, size_t n, float lr, float beta1,
                                    float beta2, float eps, float* loss, float* dgrad, void* stream) {
    return guarded([&] {
        require(mv && dx && dy && loss, "null argument");
        require(n >= 1 && n <= (1u << 22), "minibatch size must be in [1, 2^22]");
        auto* m = static_cast<StopModel*>(mv);
        DeviceGuard dg(m->device);
        hipStream_t s = (hipStream_t)stream;
        size_t blocks = (n + 255) / 256;
        m->scratch.reserve((1 + blocks + 2 * n + 4 * n * (size_t)m->hidden) * 4);
        m->step += 1;
        hip_check(hg::launch_mlp_train_step((float*)m->params.p, (float*)m->m.p, (float*)m->v.p, m->step, m->in_dim,
                                            m->hidden, dx, dy, (uint32_t)n, lr, beta1, beta2, eps,
                                            (float*)m->scratch.p, dgrad, s),
                  "mlp train step");
        hip_check(hipMemcpyAsync(loss, m->scratch.p, 4, hipMemcpyDeviceToHost, s), "hipMemcpy D2H");
        hip_check(hipStreamSynchronize(s), "mlp train step");
        *loss /= (float)n;
        return HG_OK;
    });
}
