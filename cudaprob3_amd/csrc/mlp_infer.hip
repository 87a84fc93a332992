// Fused inference of the surrogate network: one launch runs all six layers over a tile of 64 or
// 128 rows, with the activations resident in LDS between layers (no HBM traffic but the
// 36-byte input and output rows).
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
// Variants: 128-row tiles with 4 wavefronts (each over all rows; the default, fastest: 7.1e8
// rows/s) or 8 (row halves x 4 feature slices), one workgroup per CU (133 KB of LDS); 64-row
// tiles with 4 wavefronts and 66.5 KB (two workgroups per CU where registers allow).  Within a
// wavefront the GELU epilogue of a 16-row block follows its MFMAs; the compiler does not
// interleave the two.
//
// Grid mode (kGrid): the rows are the points (hypothesis h, cosine ic, energy bin ib) of
// n_hyp hypotheses on one cosZ x bin grid, h-major then ic then ib; the kernel forms each
// row's inputs (cos[ic], log10e[ib], the hypothesis' 6 parameters and nu / nubar flag) itself
// and stores the outputs in the exact kernels' SoA layout, y[h][t][ic][ib].  With skip_below,
// rows with ib < skip_below[h * n_cos + ic] (bins that the exact kernel provides) are not
// stored, so the two kernels fill disjoint parts of one array, and a tile made only of such
// rows is skipped after one barrier (persistent launch, see mlp_forward_fused) -- or, with a
// list of the tiles to compute, only those tiles are launched.
#include "common.hpp"
#include "mlp.hpp"
#include "mlp_device.hpp"

#include <algorithm>

namespace cp3 {
namespace {

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

// grid mode: (hypothesis, cosine, bin) of each row of the tile, computed once per tile (one
// 64-bit division per thread) and kept in LDS for the input and output stages
template <int kTileRows>
struct TileRows {
    int h[kTileRows], ic[kTileRows], ib[kTileRows];
};

// One GELU layer over the tile: act (hi/lo) -> act (hi/lo).  kFirst: the 9 -> 256 layer, whose
// K = 32 columns already hold both terms of the inputs (one MFMA per feature block).
template <int kTileRows, int kRowGroups, bool kFirst>
__device__ __attribute__((always_inline)) inline void hidden_layer(__bf16* act, const __bf16* __restrict__ w,
                                                                   const float* __restrict__ bias, int wm, int wn,
                                                                   int lane) {
    constexpr int kMT = kTileRows / 16 / kRowGroups;       // 16-row tiles of this wavefront
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
        for (int j = 0; j < 4; j++) acc[j] = bv[j];         // the bias, then the products
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
            float g[4];
#pragma unroll
            for (int r = 0; r < 4; r++) g[r] = gelu(acc[j][r]);
            two_term_pair(g[0], g[1], out_hi[i][j].x, out_lo[i][j].x);
            two_term_pair(g[2], g[3], out_hi[i][j].y, out_lo[i][j].y);
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
template <int kTileRows, int kRowGroups, bool kGrid>
__device__ __attribute__((always_inline)) inline void output_layer(const __bf16* act, const __bf16* __restrict__ w,
                                                                   const float* __restrict__ bias, float* __restrict__ y,
                                                                   long long row0, long long n, const MlpGrid& grid,
                                                                   const TileRows<kTileRows>& tr, int wave, int lane) {
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
                const int r16 = i * 16 + l16;
                const int h = tr.h[r16], ic = tr.ic[r16], ib = tr.ib[r16];
                const long long plane = (long long)grid.n_cos * grid.n_bins;
                if (!grid.skip_below || ib >= grid.skip_below[(long long)h * grid.n_cos + ic]) {
                    float* o = y + (long long)h * kMlpOut * plane + (long long)ic * grid.n_bins + ib;
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

// One tile: rows row0 .. row0 + kTileRows - 1 through all layers.
template <int kTileRows, int kRowGroups, bool kGrid>
__device__ __attribute__((always_inline)) inline void fused_tile(const FusedArgs& a, __bf16* act,
                                                                 TileRows<kTileRows>& tr, long long row0, int wave,
                                                                 int lane) {
    if (kGrid) {
        int needed = 0;
        if (threadIdx.x < kTileRows) {
            const long long row = row0 + threadIdx.x;
            const long long plane = (long long)a.g.n_cos * a.g.n_bins;
            const long long h = row / plane;
            const int rr = int(row - h * plane);
            const int ic = rr / a.g.n_bins, ib = rr - ic * a.g.n_bins;
            tr.h[threadIdx.x] = int(h);
            tr.ic[threadIdx.x] = ic;
            tr.ib[threadIdx.x] = ib;
            needed = row < a.n && (!a.g.skip_below || ib >= a.g.skip_below[h * a.g.n_cos + ic]);
        }
        if (!__syncthreads_or(needed)) return;      // the whole workgroup skips the tile together
    }

    // scaled inputs as two terms: hi in columns 0..8, lo in kMlpInLo.., zeros up to column 31
#pragma unroll 4
    for (int e = threadIdx.x; e < kTileRows * kMlpInPad; e += 256 * kRowGroups) {    // blockDim.x = 256 * kRowGroups
        const int r = e / kMlpInPad, c = e % kMlpInPad;
        const int src = c < kMlpIn ? c : c - kMlpInLo;
        float v = 0.0f;
        if (row0 + r < a.n && src < kMlpIn) {
            float x;
            if (kGrid) {
                x = src == 0   ? a.g.cos[tr.ic[r]]
                    : src == 1 ? a.g.log10e[tr.ib[r]]
                               : a.g.hyp[(long long)tr.h[r] * (kMlpIn - 2) + src - 2];
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
    hidden_layer<kTileRows, kRowGroups, true>(act, a.w[0], a.b[0], wm, wn, lane);
#pragma unroll 1
    for (int l = 1; l < kMlpHiddenLayers; l++)
        hidden_layer<kTileRows, kRowGroups, false>(act, a.w[l], a.b[l], wm, wn, lane);
    output_layer<kTileRows, kRowGroups, kGrid>(act, a.w[kMlpLayers - 1], a.b[kMlpLayers - 1], a.y, row0, a.n, a.g, tr,
                                               wave, lane);
}

// Workgroup b takes tiles b, b + gridDim, b + 2 gridDim, ...: one tile each for an ordinary
// launch (gridDim = tiles); in grid mode with skipped bins the launch is persistent (one
// workgroup per resident slot, minus one so the stride is coprime with the tiles of a cosine
// row) and unneeded tiles cost one barrier, with no dependence on how the hardware deals
// workgroups to CUs (a per-row skip pattern in a one-tile-per-workgroup launch leaves CUs that
// receive only skipped tiles idle while others queue).
template <int kTileRows, int kRowGroups, bool kGrid, bool kPersistent>
__global__ __launch_bounds__(256 * kRowGroups) void mlp_forward_fused(FusedArgs a) {
    __shared__ __bf16 act[kTileRows * kActStride];
    __shared__ TileRows<kTileRows> tr;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (!kPersistent) {
        const long long tile = kGrid && a.g.tiles ? a.g.tiles[blockIdx.x] : (long long)blockIdx.x;
        fused_tile<kTileRows, kRowGroups, kGrid>(a, act, tr, tile * kTileRows, wave, lane);
        return;
    }
    const long long tiles = (a.n + kTileRows - 1) / kTileRows;
    for (long long tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        if (tile != (long long)blockIdx.x) __syncthreads();     // the previous tile's LDS reads are done
        fused_tile<kTileRows, kRowGroups, kGrid>(a, act, tr, tile * kTileRows, wave, lane);
    }
}

}  // namespace

hipError_t launch_mlp_forward_fused(const float* x, long long n, float* y, const __bf16* const* w,
                                    const float* const* b, const MlpScaling& sc, int variant, const MlpGrid* grid,
                                    int cu_count, long long list_len, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    FusedArgs a{x, n, y, {}, {}, sc, grid ? *grid : MlpGrid{}};
    for (int l = 0; l < kMlpLayers; l++) {
        a.w[l] = w[l];
        a.b[l] = b[l];
    }
    if (grid && (grid->n_cos <= 0 || grid->n_bins <= 0)) return hipErrorInvalidValue;
    if (grid && grid->tiles) {
        // listed 128-row tiles, one workgroup each (no skipped workgroups)
        if (variant == kFusedRows64) variant = kFusedRows128Waves4;
        if (list_len <= 0) return list_len == 0 ? hipSuccess : hipErrorInvalidValue;
        if (list_len > 0x7fffffffLL) return hipErrorInvalidValue;
        if (variant == kFusedRows128Waves8)
            hipLaunchKernelGGL((mlp_forward_fused<128, 2, true, false>), dim3(unsigned(list_len)), dim3(512), 0, s, a);
        else
            hipLaunchKernelGGL((mlp_forward_fused<128, 1, true, false>), dim3(unsigned(list_len)), dim3(256), 0, s, a);
        return hipGetLastError();
    }
    const int tile = variant == kFusedRows64 ? 64 : 128;
    const long long blocks = (n + tile - 1) / tile;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    // persistent launch (grid mode with skipped bins): resident workgroups of this kernel, minus one
    auto resident = [&](const void* kernel, int threads) -> long long {
        int per_cu = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu < 1)
            per_cu = 1;
        return std::max(1LL, (long long)per_cu * cu_count - 1);
    };
#define CP3_FUSED(T, G)                                                                                           \
    do {                                                                                                          \
        if (grid && grid->skip_below) {                                                                           \
            const void* k = reinterpret_cast<const void*>(&mlp_forward_fused<T, G, true, true>);                 \
            const long long launch = std::min(blocks, resident(k, 256 * G));                                     \
            hipLaunchKernelGGL((mlp_forward_fused<T, G, true, true>), dim3(unsigned(launch)), dim3(256 * G), 0, s, a); \
        } else if (grid) {                                                                                        \
            hipLaunchKernelGGL((mlp_forward_fused<T, G, true, false>), dim3(unsigned(blocks)), dim3(256 * G), 0, s, a); \
        } else {                                                                                                  \
            hipLaunchKernelGGL((mlp_forward_fused<T, G, false, false>), dim3(unsigned(blocks)), dim3(256 * G), 0, s, a); \
        }                                                                                                         \
    } while (0)
    switch (variant) {
        case kFusedRows128Waves4: CP3_FUSED(128, 1); break;
        case kFusedRows128Waves8: CP3_FUSED(128, 2); break;
        case kFusedRows64: CP3_FUSED(64, 1); break;
        default: return hipErrorInvalidValue;
    }
#undef CP3_FUSED
    return hipGetLastError();
}

}  // namespace cp3
