// Host-side helpers shared by the translation units of libmmhip: the
// thread-local error message, HIP error checks, and the per-device state
// (sqrt / 1/sqrt tables of n < 100000, LDS limit, CU count).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mm_hermite.h"

namespace mm {

inline thread_local std::string g_err = "no error";

inline int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return fail(MM_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define LAUNCH_CHECK(what)                                                                      \
    do {                                                                                        \
        hipError_t e_ = hipGetLastError();                                                      \
        if (e_ != hipSuccess) return fail(MM_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e_)); \
    } while (0)

struct DevState {
    double *sq = nullptr, *rsq = nullptr;
    int max_lds = 65536;
    int num_cu = 1;
};

inline std::mutex g_mu;
inline std::map<int, DevState> g_dev;

// State of the current device, created on first use (tables uploaded once).
inline int device_state(DevState **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_dev.find(dev);
    if (it != g_dev.end()) {
        *out = &it->second;
        return MM_OK;
    }
    DevState st;
    std::vector<double> sq(MM_MAX_CUTOFF), rsq(MM_MAX_CUTOFF);
    for (int n = 0; n < MM_MAX_CUTOFF; ++n) {
        sq[n] = std::sqrt(double(n));  // correctly rounded, as np.sqrt (core.py:22)
        rsq[n] = n ? 1.0 / sq[n] : 0.0;
    }
    HIP_TRY(hipMalloc(&st.sq, sizeof(double) * MM_MAX_CUTOFF));
    HIP_TRY(hipMalloc(&st.rsq, sizeof(double) * MM_MAX_CUTOFF));
    HIP_TRY(hipMemcpy(st.sq, sq.data(), sizeof(double) * MM_MAX_CUTOFF, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(st.rsq, rsq.data(), sizeof(double) * MM_MAX_CUTOFF, hipMemcpyHostToDevice));
    HIP_TRY(hipDeviceGetAttribute(&st.max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    HIP_TRY(hipDeviceGetAttribute(&st.num_cu, hipDeviceAttributeMultiprocessorCount, dev));
    st.max_lds -= 256;  // headroom for the kernels' static LDS
    auto ins = g_dev.emplace(dev, st);
    *out = &ins.first->second;
    return MM_OK;
}

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace mm
