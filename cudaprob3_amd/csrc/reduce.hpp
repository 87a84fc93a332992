// Deterministic FP64 sums across a wavefront and a workgroup (fixed combination order), shared
// by the reduction kernels and the likelihood variant of the mixed pair kernel.
#pragma once

#include "common.hpp"

namespace cp3 {

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Sum of v over the workgroup, valid in thread 0; lds holds one double per wavefront.
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

}  // namespace cp3
