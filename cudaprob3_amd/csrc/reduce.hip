// Device-side comparison of a result grid with reference data: sum of squared differences
// in FP64, so a likelihood evaluation (e.g. a Markov-chain step) never copies the grid to
// the host.  Two passes for a deterministic sum: per-workgroup partials, then one
// workgroup adds the partials in a fixed order.
#include "common.hpp"

namespace cp3 {
namespace {

constexpr int kReduceBlocks = 1024;

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Sum of v over the workgroup, valid in thread 0.
__device__ inline double block_sum(double v, double* lds) {
    v = wave_sum(v);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < int(blockDim.x) / kWave; w++) s += lds[w];
    return s;
}

__global__ __launch_bounds__(kBlock) void sqdiff_partials(const float* __restrict__ a, const float* __restrict__ b,
                                                          std::size_t n, double* __restrict__ partials) {
    __shared__ double lds[kBlock / kWave];
    double acc = 0.0;
    const std::size_t stride = std::size_t(gridDim.x) * blockDim.x;
    const std::size_t n4 = n / 4;
    const float4* a4 = reinterpret_cast<const float4*>(a);
    const float4* b4 = reinterpret_cast<const float4*>(b);
    for (std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n4; i += stride) {
        const float4 x = a4[i], y = b4[i];
        const double d0 = double(x.x) - y.x, d1 = double(x.y) - y.y, d2 = double(x.z) - y.z, d3 = double(x.w) - y.w;
        acc += d0 * d0 + d1 * d1 + d2 * d2 + d3 * d3;
    }
    for (std::size_t i = 4 * n4 + std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double d = double(a[i]) - b[i];
        acc += d * d;
    }
    const double s = block_sum(acc, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void sum_partials(const double* __restrict__ partials, int n,
                                                       double* __restrict__ out) {
    __shared__ double lds[kBlock / kWave];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partials[i];
    const double s = block_sum(acc, lds);
    if (threadIdx.x == 0) out[0] = s;
}

}  // namespace

int reduce_partials_needed() { return kReduceBlocks; }

// a, b: device float arrays of n elements (16-byte aligned); partials: kReduceBlocks + 1
// doubles of device scratch; the sum lands in partials[kReduceBlocks].
hipError_t launch_sqdiff_sum(const float* a, const float* b, std::size_t n, double* partials, hipStream_t s) {
    hipLaunchKernelGGL(sqdiff_partials, dim3(kReduceBlocks), dim3(kBlock), 0, s, a, b, n, partials);
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(kBlock), 0, s, partials, kReduceBlocks, partials + kReduceBlocks);
    return hipGetLastError();
}

}  // namespace cp3
