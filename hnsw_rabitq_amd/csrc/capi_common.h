// Helpers shared by the C ABI translation units (error capture, device buffers).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/hnswgpu.h"
#include "common.h"

#define HG_API extern "C" __attribute__((visibility("default")))

namespace hgapi {

inline thread_local std::string g_err;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(HG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
inline int guarded(F f) {
    try {
        return f();
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "Not enough memory";
        return HG_ERR_RUNTIME;
    } catch (const std::exception& e) {
        g_err = e.what();
        return HG_ERR_RUNTIME;
    }
}

inline void require(bool c, const std::string& msg) {
    if (!c) throw Error(HG_ERR_INVALID, msg);
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void reserve(size_t b) {
        if (b <= bytes) return;
        if (p) hip_check(hipFree(p), "hipFree");
        p = nullptr;
        bytes = 0;
        hip_check(hipMalloc(&p, b), "hipMalloc");
        bytes = b;
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        hip_check(hipGetDevice(&prev), "hipGetDevice");
        if (prev != dev) hip_check(hipSetDevice(dev), "hipSetDevice");
    }
    ~DeviceGuard() {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
    }
};

// ------------------------------------------------------------------------------------------
// Stop model (learned early termination) device state.
// ------------------------------------------------------------------------------------------
struct StopModel {
    int device = 0;
    int in_dim = 0, hidden = 0;
    size_t nparam = 0;
    DevBuf params, m, v, mean, inv_std, scratch;
    int step = 0;
    int period = 16, min_hops = 16;
    float threshold = 0.9f;
    // extension: at the end of a search, a query whose confidence is below ext_threshold is
    // answered again at ext_ef (0 = no extension)
    int ext_ef = 0;
    float ext_threshold = 0.5f;
};

inline size_t mlp_param_count(int in, int h) { return (size_t)h * in + h + (size_t)h * h + h + h + 1; }

inline hg::DevStopModel dev_stop(StopModel* m) {
    hg::DevStopModel d{};
    d.params = (const float*)m->params.p;
    d.mean = (const float*)m->mean.p;
    d.inv_std = (const float*)m->inv_std.p;
    d.in_dim = m->in_dim;
    d.hidden = m->hidden;
    d.period = m->period;
    d.min_hops = m->min_hops;
    float t = std::min(std::max(m->threshold, 1e-6f), 1.0f - 1e-6f);
    d.logit_threshold = std::log(t / (1.0f - t));
    float te = std::min(std::max(m->ext_threshold, 1e-6f), 1.0f - 1e-6f);
    d.ext_logit_threshold = m->ext_ef > 0 ? std::log(te / (1.0f - te)) : -INFINITY;
    return d;
}

}  // namespace hgapi
