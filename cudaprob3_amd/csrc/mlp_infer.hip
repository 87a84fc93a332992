// Fused inference of the surrogate network: one launch runs all six layers over a tile of 128
// rows, with the activations resident in LDS between layers (no HBM traffic but the 36-byte
// input and output rows).
//
// LDS holds the tile's activations as two-term bf16 (hi in columns 0..255, lo in 256..511;
// before the first layer, the scaled inputs' hi / lo in columns 0..31), rows padded to 520
// bf16 so that the 16 rows of an MFMA operand read fall on different banks.
//
// Each wavefront owns 64 output features of every hidden layer and keeps their bf16 weights in
// registers for the whole layer (4 feature blocks x 8 k-steps x 8 bf16 = 128 VGPRs).  The
// product is taken transposed, C^T = W A^T with v_mfma_f32_16x16x32_bf16 (operand A = 16
// weight rows, operand B = 16 activation rows), so a lane's four results are four consecutive
// features of one row and go back to LDS as one 8-byte store per term.  The hi and lo terms
// of an activation row meet the same weight registers (two MFMAs per k-step).  Results wait in
// registers until every wavefront has finished reading the layer's inputs (workgroup barrier),
// then overwrite them in place.
//
// kRowGroups = 1: 4 wavefronts, each over all 128 rows; kRowGroups = 2: 8 wavefronts, row
// halves x 4 feature slices.
//
// Grid mode (kGrid): the rows are the points (hypothesis h, cosine ic, energy bin ib) of
// n_hyp hypotheses on one cosZ x bin grid, h-major then ic then ib; the kernel forms each
// row's inputs (cos[ic], log10e[ib], the hypothesis' 6 parameters and nu / nubar flag) itself
// and stores the outputs in the exact kernels' SoA layout, y[h][t][ic][ib].  With skip_below,
// rows with ib < skip_below[h * n_cos + ic] (bins that the exact kernel provides) are not
// stored, so the two kernels can fill one array concurrently, and a tile made only of such
// rows exits before any work.
#include "common.hpp"
#include "mlp.hpp"
#include "mlp_device.hpp"

namespace cp3 {
namespace {

constexpr int kTileRows = 128;
constexpr int kActStride = kMlpActLd + 8;

struct FusedArgs {
    const float* x;                    // [n][9] raw inputs (row mode)
    long long n;
    float* y;                          // [n][9] (row mode) or [n_hyp][9][n_cos][n_bins] (grid mode)
    const __bf16* w[kMlpLayers];       // bf16 weights, two-term layout (mlp.hpp)
    const float* b[kMlpLayers];
    MlpScaling sc;
    MlpGrid g;                         // grid mode
};

// grid mode: hypothesis and in-plane index of a row
struct GridRow {
    long long h;
    int rr;      // ic * n_bins + ib
};

__device__ inline GridRow grid_row(const MlpGrid& g, long long row) {
    const long long plane = (long long)g.n_cos * g.n_bins;
    const long long h = row / plane;
    return GridRow{h, int(row - h * plane)};
}

__device__ inline unsigned bf16_pair(__bf16 lo_half, __bf16 hi_half) {
    return unsigned(__builtin_bit_cast(unsigned short, lo_half)) |
           (unsigned(__builtin_bit_cast(unsigned short, hi_half)) << 16);
}

// One GELU layer over the tile: act (hi/lo) -> act (hi/lo).  kFirst: the 9 -> 256 layer, whose
// K = 32 columns already hold both terms of the inputs (one MFMA per feature block).
template <int kRowGroups, bool kFirst>
__device__ inline void hidden_layer(__bf16* act, const __bf16* __restrict__ w, const float* __restrict__ bias,
                                    int wm, int wn, int lane) {
    constexpr int kMT = 8 / kRowGroups;                    // 16-row tiles of this wavefront
    constexpr int kKS = kFirst ? 1 : kMlpHidden / 32;      // k-steps over one term
    constexpr int kLd = kFirst ? kMlpInPad : kMlpActLd;    // row length of w
    const int l16 = lane % 16, q = lane / 16;

    bf16x8 wf[4][kKS];
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int kk = 0; kk < kKS; kk++)
            wf[j][kk] = *reinterpret_cast<const bf16x8*>(w + std::size_t(wn * 64 + j * 16 + l16) * kLd + kk * 32 + 8 * q);
    f32x4 bv[4];
#pragma unroll
    for (int j = 0; j < 4; j++) bv[j] = *reinterpret_cast<const f32x4*>(bias + wn * 64 + j * 16 + 4 * q);

    uint2 out_hi[kMT][4], out_lo[kMT][4];
#pragma unroll
    for (int i = 0; i < kMT; i++) {
        const __bf16* arow = act + (wm * kMT * 16 + i * 16 + l16) * kActStride + 8 * q;
        f32x4 acc[4];
#pragma unroll
        for (int j = 0; j < 4; j++) acc[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kk = 0; kk < kKS; kk++) {
            const bf16x8 ah = *reinterpret_cast<const bf16x8*>(arow + kk * 32);
#pragma unroll
            for (int j = 0; j < 4; j++) acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[j][kk], ah, acc[j], 0, 0, 0);
            if (!kFirst) {
                const bf16x8 al = *reinterpret_cast<const bf16x8*>(arow + kMlpHidden + kk * 32);
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[j][kk], al, acc[j], 0, 0, 0);
            }
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            __bf16 h[4], lo[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float g = gelu(acc[j][r] + bv[j][r]);
                h[r] = __bf16(g);
                lo[r] = __bf16(g - float(h[r]));
            }
            out_hi[i][j] = uint2{bf16_pair(h[0], h[1]), bf16_pair(h[2], h[3])};
            out_lo[i][j] = uint2{bf16_pair(lo[0], lo[1]), bf16_pair(lo[2], lo[3])};
        }
    }
    __syncthreads();    // every wavefront has read this layer's inputs
#pragma unroll
    for (int i = 0; i < kMT; i++) {
        __bf16* orow = act + (wm * kMT * 16 + i * 16 + l16) * kActStride + wn * 64 + 4 * q;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            *reinterpret_cast<uint2*>(orow + j * 16) = out_hi[i][j];
            *reinterpret_cast<uint2*>(orow + kMlpHidden + j * 16) = out_lo[i][j];
        }
    }
    __syncthreads();
}

// 256 -> 9 linear layer (16 computed features, 9 stored); the wavefronts take 16-row tiles in turn.
template <int kRowGroups, bool kGrid>
__device__ inline void output_layer(const __bf16* act, const __bf16* __restrict__ w, const float* __restrict__ bias,
                                    float* __restrict__ y, long long row0, long long n, const MlpGrid& grid, int wave,
                                    int lane) {
    constexpr int kWaves = 4 * kRowGroups;
    constexpr int kKS = kMlpHidden / 32;
    const int l16 = lane % 16, q = lane / 16;
    bf16x8 wf[kKS];
#pragma unroll
    for (int kk = 0; kk < kKS; kk++)
        wf[kk] = *reinterpret_cast<const bf16x8*>(w + std::size_t(l16) * kMlpActLd + kk * 32 + 8 * q);
    const f32x4 bv = *reinterpret_cast<const f32x4*>(bias + 4 * q);
    for (int i = wave; i < kTileRows / 16; i += kWaves) {
        const __bf16* arow = act + (i * 16 + l16) * kActStride + 8 * q;
        f32x4 acc{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kk = 0; kk < kKS; kk++) {
            const bf16x8 ah = *reinterpret_cast<const bf16x8*>(arow + kk * 32);
            const bf16x8 al = *reinterpret_cast<const bf16x8*>(arow + kMlpHidden + kk * 32);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[kk], ah, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[kk], al, acc, 0, 0, 0);
        }
        const long long row = row0 + i * 16 + l16;         // lane holds features 4q .. 4q+3 of this row
        if (row < n) {
            if (kGrid) {
                const GridRow gr = grid_row(grid, row);
                const long long plane = (long long)grid.n_cos * grid.n_bins;
                const int ic = gr.rr / grid.n_bins, ib = gr.rr % grid.n_bins;
                if (!grid.skip_below || ib >= grid.skip_below[gr.h * grid.n_cos + ic]) {
                    float* o = y + gr.h * kMlpOut * plane + gr.rr;
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (4 * q + r < kMlpOut) o[(4 * q + r) * plane] = acc[r] + bv[r];
                }
            } else {
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (4 * q + r < kMlpOut) y[row * kMlpOut + 4 * q + r] = acc[r] + bv[r];
            }
        }
    }
}

template <int kRowGroups, bool kGrid>
__global__ __launch_bounds__(256 * kRowGroups) void mlp_forward_fused(FusedArgs a) {
    __shared__ __bf16 act[kTileRows * kActStride];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const long long row0 = (long long)blockIdx.x * kTileRows;

    if (kGrid && a.g.skip_below) {
        int needed = 0;
        if (threadIdx.x < kTileRows && row0 + threadIdx.x < a.n) {
            const GridRow gr = grid_row(a.g, row0 + threadIdx.x);
            const int ic = gr.rr / a.g.n_bins, ib = gr.rr % a.g.n_bins;
            needed = ib >= a.g.skip_below[gr.h * a.g.n_cos + ic];
        }
        if (!__syncthreads_or(needed)) return;      // the whole workgroup leaves together
    }

    // scaled inputs as two terms: hi in columns 0..8, lo in kMlpInLo.., zeros up to column 31
    for (int e = threadIdx.x; e < kTileRows * kMlpInPad; e += blockDim.x) {
        const int r = e / kMlpInPad, c = e % kMlpInPad;
        const int src = c < kMlpIn ? c : c - kMlpInLo;
        float v = 0.0f;
        if (row0 + r < a.n && src < kMlpIn) {
            float x;
            if (kGrid) {
                const GridRow gr = grid_row(a.g, row0 + r);
                x = src == 0   ? a.g.cos[gr.rr / a.g.n_bins]
                    : src == 1 ? a.g.log10e[gr.rr % a.g.n_bins]
                               : a.g.hyp[gr.h * (kMlpIn - 2) + src - 2];
            } else {
                x = a.x[(row0 + r) * kMlpIn + src];
            }
            v = (x - a.sc.shift[src]) * a.sc.scale[src];
        }
        const __bf16 hi = __bf16(v);
        act[r * kActStride + c] = c < kMlpIn ? hi : __bf16(v - float(hi));
    }
    __syncthreads();

    const int wm = wave / 4, wn = wave % 4;
    hidden_layer<kRowGroups, true>(act, a.w[0], a.b[0], wm, wn, lane);
#pragma unroll 1
    for (int l = 1; l < kMlpHiddenLayers; l++) hidden_layer<kRowGroups, false>(act, a.w[l], a.b[l], wm, wn, lane);
    output_layer<kRowGroups, kGrid>(act, a.w[kMlpLayers - 1], a.b[kMlpLayers - 1], a.y, row0, a.n, a.g, wave, lane);
}

}  // namespace

hipError_t launch_mlp_forward_fused(const float* x, long long n, float* y, const __bf16* const* w,
                                    const float* const* b, const MlpScaling& sc, int row_groups, const MlpGrid* grid,
                                    hipStream_t s) {
    if (n <= 0) return hipSuccess;
    FusedArgs a{x, n, y, {}, {}, sc, grid ? *grid : MlpGrid{}};
    for (int l = 0; l < kMlpLayers; l++) {
        a.w[l] = w[l];
        a.b[l] = b[l];
    }
    const long long blocks = (n + kTileRows - 1) / kTileRows;
    if (blocks > 0x7fffffffLL || (grid && (grid->n_cos <= 0 || grid->n_bins <= 0))) return hipErrorInvalidValue;
#define CP3_FUSED(G, GRID) \
    hipLaunchKernelGGL((mlp_forward_fused<G, GRID>), dim3(unsigned(blocks)), dim3(256 * G), 0, s, a)
    if (row_groups == 1) {
        if (grid) CP3_FUSED(1, true); else CP3_FUSED(1, false);
    } else if (row_groups == 2) {
        if (grid) CP3_FUSED(2, true); else CP3_FUSED(2, false);
    } else {
        return hipErrorInvalidValue;
    }
#undef CP3_FUSED
    return hipGetLastError();
}

}  // namespace cp3
