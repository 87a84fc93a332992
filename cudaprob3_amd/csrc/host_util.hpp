// Host-side helpers shared by the C ABI translation units.
#pragma once

#include "../../include/cp3hip.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

namespace cp3host {

// Message of the last failing call on this thread (cp3h_last_error).
inline thread_local std::string g_error;

inline int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

#define CP3H_HIP(call)                                                                           \
    do {                                                                                         \
        hipError_t err_ = (call);                                                                \
        if (err_ != hipSuccess)                                                                  \
            return ::cp3host::fail(CP3H_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(err_)); \
    } while (0)

// Vacuum cubic roots matched to DM(i,0) (physics.hpp:198-253), in the caller's precision:
// order[i] is the matter root that continues vacuum mass i.
template <typename T>
void vacuum_order(const T* DM, int order[3]) {
    const T a = DM[1] + DM[2];
    const T b = DM[1] * DM[2];
    const T g = 0.0;
    T t = a * a - T(3.0) * b;
    T arg = (T(2.0) * a * a * a - T(9.0) * a * b + T(27.0) * g) / (T(2.0) * std::sqrt(t * t * t));
    if (std::fabs(arg) > T(1.0)) arg = arg / std::fabs(arg);
    const T th0 = std::acos(arg) / T(3.0);
    const T th[3] = {th0, T(th0 - (2.0 * M_PI / 3.0)), T(th0 + (2.0 * M_PI / 3.0))};
    T root[3];
    for (int k = 0; k < 3; k++) {
        root[k] = T(-(2.0 / 3.0) * std::sqrt(t));
        root[k] *= std::cos(th[k]);
    }
    t = DM[0] - a / T(3.0);
    for (int k = 0; k < 3; k++) root[k] += t;
    for (int i = 0; i < 3; i++) {
        T best = std::fabs(DM[i * 3] - root[0]);
        int k = 0;
        for (int j = 1; j < 3; j++) {
            const T d = std::fabs(DM[i * 3] - root[j]);
            if (d < best) { k = j; best = d; }
        }
        order[i] = k;
    }
}

}  // namespace cp3host
