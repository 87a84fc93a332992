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

// Two-term activations.  The forward pass carries every activation (and every scaled input)
// as a pair of bf16 values, hi = bf16(a) and lo = bf16(a - hi), whose sum holds a to about
// 16 significant bits; a product with a bf16 weight row then takes two MFMA k-steps against
// the same weights.  The layouts: scaled inputs [M][kMlpInPad] with hi in columns 0..8 and lo
// in columns kMlpInLo..kMlpInLo+8; hidden activations [M][kMlpActLd], hi in 0..255, lo in
// 256..511; the bf16 weights repeat each row's K values at the lo offset.  Single-term bf16
// activations alone put errors of order 1e-3..1e-2 on the outputs.
constexpr int kMlpInLo = kMlpIn;
constexpr int kMlpActLd = 2 * kMlpHidden;

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
    int lo_off;        // kEpiBiasGelu: column offset of the lo terms in yout (0: hi only)
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
// wb: bf16 weights [N][ld_wb], each row's K values at column 0 and again at column `dup`
hipError_t launch_adam_layer(float* w, float* m1, float* m2, const float* g, int N, int K, int n_valid, int k_valid,
                             const AdamStep& a, __bf16* wb, int ld_wb, int dup, __bf16* wtb, hipStream_t s);
// Grid-mode inference: rows (h, ic, ib) of n_hyp hypotheses on one cosZ x energy-bin grid.
struct MlpGrid {
    const float* cos;         // [n_cos]
    const float* log10e;      // [n_bins] log10 of the bin centres / GeV
    const float* hyp;         // [n_hyp][7]: theta12, theta13, theta23, deltaCP, dm21^2, dm32^2, nubar flag
    const int* skip_below;    // optional [n_hyp][n_cos]: rows with ib below it may be left unwritten
    const long long* tiles;   // optional: the 128-row tiles to compute (tile t = rows 128 t ..), one per workgroup
    int n_cos, n_bins;
};

// all layers of inference in one launch (mlp_infer.hip), tile shape by variant;
// grid: null for row mode (x [n][9] -> y [n][9]), else grid mode (y [n_hyp][9][n_cos][n_bins])
enum FusedVariant { kFusedRows128Waves4 = 1, kFusedRows128Waves8 = 2, kFusedRows64 = 3 };
hipError_t launch_mlp_forward_fused(const float* x, long long n, float* y, const __bf16* const* w,
                                    const float* const* b, const MlpScaling& sc, int variant, const MlpGrid* grid,
                                    int cu_count, long long list_len, hipStream_t s);
hipError_t launch_adam_vec(float* w, float* m1, float* m2, const float* g, int n, int n_valid, const AdamStep& a,
                           hipStream_t s);

}  // namespace cp3
