// Device-side comparison of a result grid with reference data: sum of squared differences
// in FP64, so a likelihood evaluation (e.g. a Markov-chain step) never copies the grid to
// the host.  Two passes for a deterministic sum: per-workgroup partials, then one
// workgroup adds the partials in a fixed order.
#include "common.hpp"
#include "reduce.hpp"

namespace cp3 {
namespace {

constexpr int kReduceBlocks = 1024;

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

// out[blockIdx.x] = sum of in[i] over i = blockIdx.x * blockDim.x + threadIdx.x (mod the grid)
__global__ __launch_bounds__(kBlock) void sum_partials(const double* __restrict__ in, long long n,
                                                       double* __restrict__ out) {
    __shared__ double lds[kBlock / kWave];
    double acc = 0.0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) acc += in[i];
    const double s = block_sum(acc, lds);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

}  // namespace

int reduce_partials_needed() { return kReduceBlocks; }

// a, b: device float arrays of n elements (16-byte aligned); partials: kReduceBlocks + 1
// doubles of device scratch; the sum lands in partials[kReduceBlocks].
hipError_t launch_sqdiff_sum(const float* a, const float* b, std::size_t n, double* partials, hipStream_t s) {
    hipLaunchKernelGGL(sqdiff_partials, dim3(kReduceBlocks), dim3(kBlock), 0, s, a, b, n, partials);
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(kBlock), 0, s, partials, (long long)kReduceBlocks,
                       partials + kReduceBlocks);
    return hipGetLastError();
}

// sum of n doubles (any n >= 0) in two fixed-order passes; scratch: kReduceBlocks + 1 doubles,
// the sum lands in scratch[kReduceBlocks].
hipError_t launch_sum_f64(const double* in, long long n, double* scratch, hipStream_t s) {
    hipLaunchKernelGGL(sum_partials, dim3(kReduceBlocks), dim3(kBlock), 0, s, in, n, scratch);
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(kBlock), 0, s, scratch, (long long)kReduceBlocks,
                       scratch + kReduceBlocks);
    return hipGetLastError();
}

}  // namespace cp3
