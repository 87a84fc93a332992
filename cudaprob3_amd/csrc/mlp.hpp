// Declarations shared by the surrogate network's kernels (mlp.hip) and its C ABI
// (cp3hip_mlp.cpp).
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <cstddef>

namespace cp3 {

// Network shape: 9 inputs -> 5 hidden layers of 256 (GELU) -> 9 outputs.  The first layer's
// K and the output layer's N are padded to 32 with zero weights (MFMA granularity).
constexpr int kMlpIn = 9, kMlpInPad = 32;
constexpr int kMlpHidden = 256, kMlpHiddenLayers = 5;
constexpr int kMlpOut = 9, kMlpOutPad = 32;
constexpr int kMlpLayers = kMlpHiddenLayers + 1;

// Epilogues of the NT GEMM.
enum GemmEpilogue {
    kEpiBiasGelu = 0,   // Z = acc + b -> zout (fp32, optional); Y = gelu(Z) -> yout, ytout (bf16)
    kEpiBias = 1,       // Z = acc + b -> zout (fp32)
    kEpiGeluGrad = 2,   // dZ = acc * gelu'(zprev) -> yout, ytout (bf16)
    kEpiPartial = 3     // split-K partial sums -> zout[split][M][N] (fp32)
};

struct GemmDesc {
    const __bf16* A; int lda;
    const __bf16* B; int ldb;
    int M, N, K, k_split;
    int epilogue;
    const float* bias;
    const float* zprev; int ldz;
    float* zout; int ldzo;
    __bf16* yout; int ldy;
    __bf16* ytout; int ldyt;
    int n_valid;
};

struct MlpScaling {
    float shift[kMlpIn];
    float scale[kMlpIn];
};

struct AdamStep {
    float lr, beta1, beta2, eps, corr1, corr2;   // corr = 1 / (1 - beta^t)
};

hipError_t launch_gemm_nt(const GemmDesc& d, hipStream_t s);
hipError_t launch_featurize(const float* x, int M, const MlpScaling& sc, __bf16* f, __bf16* ft, hipStream_t s);
hipError_t launch_output_head(const float* z, int M, const float* t, float* y, __bf16* dz, __bf16* dzt,
                              float* loss_part, hipStream_t s);
int output_head_blocks(int M);
hipError_t launch_row_sums(const __bf16* a, int rows, int M, float* out, hipStream_t s);
hipError_t launch_sum_parts(const float* p, int parts, int n, float* out, hipStream_t s);
hipError_t launch_sum_floats(const float* p, int n, float* out, hipStream_t s);
hipError_t launch_adam_layer(float* w, float* m1, float* m2, const float* g, int N, int K, int n_valid, int k_valid,
                             const AdamStep& a, __bf16* wb, __bf16* wtb, hipStream_t s);
hipError_t launch_adam_vec(float* w, float* m1, float* m2, const float* g, int n, int n_valid, const AdamStep& a,
                           hipStream_t s);

}  // namespace cp3
