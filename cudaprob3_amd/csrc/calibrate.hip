// Roofline calibration: the sustained rate of plain (non-packed) vector fused multiply-adds on
// one GPU, cp3h_measure_valu_peak (include/cp3hip.h).
//
// Every lane runs kChains independent FMA chains a <- a * b + c (no dependence between chains,
// so the issue rate, not the FMA latency, bounds it), eight wavefronts per SIMD; each lane
// stores one sum at the end, so the chains cannot be removed.  The rate this reaches is the
// ceiling of a kernel whose arithmetic is FMAs issued one per instruction, beside the data
// sheet's FP32 figure (which counts packed FP32, disabled in this library).
#include "../../include/cp3hip.h"
#include "common.hpp"
#include "host_util.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

namespace {

using cp3host::fail;

constexpr int kThreads = 256;          // four wavefronts per workgroup
constexpr int kGroupsPerCu = 8;        // eight wavefronts per SIMD
constexpr int kChainsF32 = 16;
constexpr int kChainsF64 = 8;

template <typename T, int kChains>
__global__ __launch_bounds__(kThreads) void fma_stream(T* __restrict__ out, int iters, T b, T c) {
    T a[kChains];
#pragma unroll
    for (int k = 0; k < kChains; k++) a[k] = T(threadIdx.x) * T(1e-7) + T(k);
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int k = 0; k < kChains; k++) a[k] = a[k] * b + c;
    }
    T s = 0;
#pragma unroll
    for (int k = 0; k < kChains; k++) s += a[k];
    out[blockIdx.x * kThreads + threadIdx.x] = s;
}

// Best TFLOP/s (2 flops per FMA) of `reps` launches, each between its own pair of HIP events,
// after `warm` untimed launches that bring the clocks of an idle GPU up.
template <typename T, int kChains>
int time_stream(T* out, int blocks, int iters, int warm, int reps, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                float* tflops) {
    const T b = T(0.999), c = T(1e-3);
    for (int r = 0; r < warm; r++)
        hipLaunchKernelGGL((fma_stream<T, kChains>), dim3(blocks), dim3(kThreads), 0, s, out, iters, b, c);
    CP3H_HIP(hipGetLastError());
    const double flops = 2.0 * double(blocks) * kThreads * double(iters) * kChains;
    float best = 0.0f;
    for (int r = 0; r < reps; r++) {
        CP3H_HIP(hipEventRecord(e0, s));
        hipLaunchKernelGGL((fma_stream<T, kChains>), dim3(blocks), dim3(kThreads), 0, s, out, iters, b, c);
        CP3H_HIP(hipGetLastError());
        CP3H_HIP(hipEventRecord(e1, s));
        CP3H_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        CP3H_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms > 0.0f) best = std::max(best, float(flops / (double(ms) * 1e-3) / 1e12));
    }
    *tflops = best;
    return CP3H_OK;
}

}  // namespace

extern "C" int cp3h_measure_valu_peak(int device, float* fp32_tflops, float* fp64_tflops) {
    if (!fp32_tflops && !fp64_tflops) return fail(CP3H_ERR_ARGUMENT, "null output pointers");
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev == 0) return fail(CP3H_ERR_DEVICE, "No GPU found");
    if (device < 0 || device >= n_dev)
        return fail(CP3H_ERR_DEVICE, "The requested GPU Id " + std::to_string(device) + " is not available.");
    CP3H_HIP(hipSetDevice(device));
    int cus = 0;
    CP3H_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    const int blocks = std::max(1, cus) * kGroupsPerCu;
    double* out = nullptr;
    CP3H_HIP(hipMalloc(&out, sizeof(double) * std::size_t(blocks) * kThreads));
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = CP3H_OK;
    auto step = [&](hipError_t err, const char* what) {
        if (rc == CP3H_OK && err != hipSuccess) rc = fail(CP3H_ERR_HIP, std::string(what) + ": " + hipGetErrorString(err));
    };
    step(hipStreamCreate(&s), "hipStreamCreate");
    step(hipEventCreate(&e0), "hipEventCreate");
    step(hipEventCreate(&e1), "hipEventCreate");
    if (rc == CP3H_OK && fp32_tflops)
        rc = time_stream<float, kChainsF32>(reinterpret_cast<float*>(out), blocks, 8000, 40, 5, s, e0, e1, fp32_tflops);
    if (rc == CP3H_OK && fp64_tflops)
        rc = time_stream<double, kChainsF64>(out, blocks, 8000, 40, 5, s, e0, e1, fp64_tflops);
    if (e1) (void)hipEventDestroy(e1);
    if (e0) (void)hipEventDestroy(e0);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(out);
    return rc;
}
