// BF16 matrix-core kernels of the surrogate network (a 9 -> 256 x 5 -> 9 MLP with GELU).
//
// Every product is an "NT" GEMM, C[m][n] = sum_k A[m][k] B[n][k], with both operands
// K-contiguous bf16 and FP32 accumulation on v_mfma_f32_16x16x32_bf16:
//   forward          Z_l  = Y_{l-1} W_l^T            A = Y_{l-1} [M][K],  B = W_l  [N][K]
//   backward (data)  dY   = dZ_l W_l                 A = dZ_l    [M][N],  B = W_l^T [K][N]
//   backward (weight) dW  = dZ_l^T Y_{l-1}           A = dZ_l^T  [N][M],  B = Y^T  [K][M]
// so the training state keeps W and W^T, and the forward / backward epilogues also write the
// transposed activations and gradients the weight GEMM reads.
//
// Tiling: a workgroup of four wavefronts (2 x 2) owns a (32*TM) x (32*TN) block of C; each
// wavefront a (16*TM) x (16*TN) block of TM x TN MFMA tiles.  K is walked in chunks of up to
// 256: the workgroup stages the A and B rows of the chunk in LDS (rows padded by 8 bf16 so
// 16-byte fragment reads of the 16 rows of a tile hit different banks), then every
// wavefront reads its fragments from LDS.  Rows beyond M or N are zero-filled, so any M
// and any N that is a multiple of 16 work; K must be a multiple of 32.
//
// MFMA operand layout (16x16x32, lane l): A row l % 16, k = 8 (l / 16) .. +7; B column
// l % 16, same k; C column l % 16, rows 4 (l / 16) .. +3.
//
// Forward products take two-term activations (hi + lo, see mlp.hpp) as a K twice as long
// against weights stored twice per row; the backward pass uses the hi terms only (Y^T, dZ).
#include "common.hpp"
#include "mlp.hpp"
#include "mlp_device.hpp"

namespace cp3 {
namespace {

constexpr int kKChunk = 256;
constexpr int kPad = 8;

struct GemmArgs {
    const __bf16* A; int lda;      // [M][K]
    const __bf16* B; int ldb;      // [N][K]
    int M, N, K;                   // K: the range [k_begin, k_begin + K) of this launch's split
    int k_split;                   // K is cut in k_split equal parts (gridDim.z); partial results
    // epilogue operands
    const float* bias;             // [N]
    const float* zprev; int ldz;   // pre-activation for the GELU derivative, [M][N]
    float* zout; int ldzo;         // fp32 output [M][N] (pre-activation or partial sums)
    __bf16* yout; int ldy;         // bf16 output [M][N]
    __bf16* ytout; int ldyt;       // bf16 transposed output [N][M]
    int n_valid;                   // columns >= n_valid are written as zero (padding)
    int lo_off;                    // kEpiBiasGelu: lo terms at yout[row][lo_off + col] (0: none)
};

template <int TM, int TN, int kEpi>
__global__ __launch_bounds__(256) void gemm_nt(GemmArgs g) {
    constexpr int WM = 16 * TM, WN = 16 * TN;          // wavefront tile
    constexpr int BM = 2 * WM, BN = 2 * WN;            // workgroup tile
    constexpr int LD = kKChunk + kPad;
    __shared__ __bf16 As[BM * LD];
    __shared__ __bf16 Bs[BN * LD];

    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int m_blk = blockIdx.x * BM, n_blk = blockIdx.y * BN;
    const int wm = m_blk + (wave / 2) * WM, wn = n_blk + (wave % 2) * WN;
    const int k_len = g.K / g.k_split;                 // a multiple of 32 by construction
    const int k_begin = blockIdx.z * k_len;

    f32x4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    for (int k0 = 0; k0 < k_len; k0 += kKChunk) {
        const int kc = min(kKChunk, k_len - k0);       // multiple of 32
        const int vec_per_row = kc / 8;
        // stage A rows m_blk..m_blk+BM-1 and B rows n_blk..n_blk+BN-1, 16 bytes per load
        for (int v = threadIdx.x; v < (BM + BN) * vec_per_row; v += blockDim.x) {
            const int row = v / vec_per_row, c8 = (v % vec_per_row) * 8;
            bf16x8 val;
            const bool is_a = row < BM;
            const int r = is_a ? row : row - BM;
            const int grow = (is_a ? m_blk : n_blk) + r;
            const int lim = is_a ? g.M : g.N;
            if (grow < lim) {
                const __bf16* src = (is_a ? g.A + std::size_t(grow) * g.lda : g.B + std::size_t(grow) * g.ldb) +
                                    k_begin + k0 + c8;
                val = *reinterpret_cast<const bf16x8*>(src);
            } else {
#pragma unroll
                for (int e = 0; e < 8; e++) val[e] = __bf16(0.0f);
            }
            __bf16* dst = (is_a ? As : Bs) + r * LD + c8;
            *reinterpret_cast<bf16x8*>(dst) = val;
        }
        __syncthreads();
        for (int kk = 0; kk < kc; kk += 32) {
            bf16x8 a[TM], b[TN];
            const int koff = kk + 8 * (lane / 16);
#pragma unroll
            for (int i = 0; i < TM; i++)
                a[i] = *reinterpret_cast<const bf16x8*>(As + ((wave / 2) * WM + i * 16 + lane % 16) * LD + koff);
#pragma unroll
            for (int j = 0; j < TN; j++)
                b[j] = *reinterpret_cast<const bf16x8*>(Bs + ((wave % 2) * WN + j * 16 + lane % 16) * LD + koff);
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }

    // epilogue: element (row, col) of tile (i, j), register r
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = wm + i * 16 + 4 * (lane / 16) + r;
                const int col = wn + j * 16 + lane % 16;
                if (row >= g.M || col >= g.N) continue;
                float v = acc[i][j][r];
                if (kEpi == kEpiPartial) {
                    g.zout[(std::size_t(blockIdx.z) * g.M + row) * g.ldzo + col] = v;
                    continue;
                }
                if (kEpi == kEpiBiasGelu || kEpi == kEpiBias) v += g.bias[col];
                if (col >= g.n_valid) v = 0.0f;
                if (kEpi == kEpiBiasGelu) {
                    if (g.zout) g.zout[std::size_t(row) * g.ldzo + col] = v;
                    const float a = gelu(v);
                    const __bf16 y = __bf16(a);
                    g.yout[std::size_t(row) * g.ldy + col] = y;
                    if (g.lo_off) g.yout[std::size_t(row) * g.ldy + g.lo_off + col] = __bf16(a - float(y));
                    if (g.ytout) g.ytout[std::size_t(col) * g.ldyt + row] = y;
                } else if (kEpi == kEpiBias) {
                    g.zout[std::size_t(row) * g.ldzo + col] = v;
                } else {   // kEpiGeluGrad: dZ_prev = dY * gelu'(Z_prev)
                    const float d = v * gelu_grad(g.zprev[std::size_t(row) * g.ldz + col]);
                    const __bf16 db = __bf16(col < g.n_valid ? d : 0.0f);
                    g.yout[std::size_t(row) * g.ldy + col] = db;
                    if (g.ytout) g.ytout[std::size_t(col) * g.ldyt + row] = db;
                }
            }
}

// --- small kernels ---------------------------------------------------------------------------

// x[M][9] fp32 raw inputs -> v = (x - shift) * scale as two-term bf16 [M][32]: hi in columns
// 0..8, lo in kMlpInLo..kMlpInLo+8, zeros elsewhere; and the hi terms transposed, [32][M], for
// the first layer's weight gradient.
__global__ __launch_bounds__(256) void featurize(const float* __restrict__ x, int M, MlpScaling s,
                                                 __bf16* __restrict__ f, __bf16* __restrict__ ft) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M * kMlpInPad) return;
    const int m = t / kMlpInPad, c = t % kMlpInPad;
    const int src = c < kMlpIn ? c : c - kMlpInLo;
    float v = 0.0f;
    if (src >= 0 && src < kMlpIn) v = (x[std::size_t(m) * kMlpIn + src] - s.shift[src]) * s.scale[src];
    const __bf16 hi = __bf16(v);
    const __bf16 b = c < kMlpIn ? hi : __bf16(v - float(hi));
    f[t] = b;
    if (ft) ft[std::size_t(c) * M + m] = c < kMlpIn ? hi : __bf16(0.0f);
}

// output layer: y = z[:, :9] (z has kMlpOutPad columns); with targets, the loss partial per workgroup
// and dZ = 2 (z - t) / (9 M) as bf16 [M][kMlpOutPad] and [kMlpOutPad][M].
__global__ __launch_bounds__(256) void output_head(const float* __restrict__ z, int M, const float* __restrict__ t,
                                                   float* __restrict__ y, __bf16* __restrict__ dz,
                                                   __bf16* __restrict__ dzt, float* __restrict__ loss_part) {
    __shared__ float red[256];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (idx < M * kMlpOutPad) {
        const int m = idx / kMlpOutPad, c = idx % kMlpOutPad;
        const float v = z[idx];
        float d = 0.0f;
        if (c < kMlpOut) {
            if (y) y[std::size_t(m) * kMlpOut + c] = v;
            if (t) {
                const float e = v - t[std::size_t(m) * kMlpOut + c];
                l = e * e;
                d = 2.0f * e / (float(kMlpOut) * float(M));
            }
        }
        if (dz) {
            dz[idx] = __bf16(d);
            dzt[std::size_t(c) * M + m] = __bf16(d);
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (loss_part && threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// column sums of a bf16 [N][M] matrix (the bias gradient from dZ^T), one workgroup per row
__global__ __launch_bounds__(256) void row_sums(const __bf16* __restrict__ a, int M, float* __restrict__ out) {
    __shared__ float red[256];
    float s = 0.0f;
    for (int m = threadIdx.x; m < M; m += blockDim.x) s += float(a[std::size_t(blockIdx.x) * M + m]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// sum of `parts` partial matrices [parts][n] -> out[n], fixed order
__global__ __launch_bounds__(256) void sum_parts(const float* __restrict__ p, int parts, int n, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int k = 0; k < parts; k++) s += p[std::size_t(k) * n + i];
    out[i] = s;
}

__global__ __launch_bounds__(256) void sum_floats(const float* __restrict__ p, int n, float* __restrict__ out) {
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// Adam on one layer's padded fp32 weights [N][K] (+ bias [N]); gradients of padding are zero.
// Writes the bf16 copies: W [N][ld_wb] with each row's K values at column 0 and again at
// column `dup` (the weights seen by the lo terms), and W^T [K][N].
__global__ __launch_bounds__(256) void adam_layer(float* __restrict__ w, float* __restrict__ m1, float* __restrict__ m2,
                                                  const float* __restrict__ g, int N, int K, int n_valid, int k_valid,
                                                  AdamStep a, __bf16* __restrict__ wb, int ld_wb, int dup,
                                                  __bf16* __restrict__ wtb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * K) return;
    const int n = i / K, k = i % K;
    float wi = w[i];
    if (n < n_valid && k < k_valid) {
        const float gi = g[i];
        const float a1 = a.beta1 * m1[i] + (1.0f - a.beta1) * gi;
        const float a2 = a.beta2 * m2[i] + (1.0f - a.beta2) * gi * gi;
        m1[i] = a1;
        m2[i] = a2;
        wi -= a.lr * (a1 * a.corr1) / (sqrtf(a2 * a.corr2) + a.eps);
        w[i] = wi;
    }
    const __bf16 b = __bf16(wi);
    if (wb) {
        if (k < dup) wb[std::size_t(n) * ld_wb + k] = b;
        if (dup + k < ld_wb) wb[std::size_t(n) * ld_wb + dup + k] = b;
    }
    if (wtb) wtb[std::size_t(k) * N + n] = b;
}

__global__ __launch_bounds__(256) void adam_vec(float* __restrict__ w, float* __restrict__ m1, float* __restrict__ m2,
                                                const float* __restrict__ g, int n, int n_valid, AdamStep a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || i >= n_valid) return;
    const float gi = g[i];
    const float a1 = a.beta1 * m1[i] + (1.0f - a.beta1) * gi;
    const float a2 = a.beta2 * m2[i] + (1.0f - a.beta2) * gi * gi;
    m1[i] = a1;
    m2[i] = a2;
    w[i] -= a.lr * (a1 * a.corr1) / (sqrtf(a2 * a.corr2) + a.eps);
}

unsigned blocks_for(long long n) { return unsigned((n + 255) / 256); }

}  // namespace

hipError_t launch_gemm_nt(const GemmDesc& d, hipStream_t s) {
    if (d.K % 32 != 0 || d.N % 16 != 0 || d.k_split < 1 || (d.K / d.k_split) % 32 != 0 || d.K % d.k_split != 0)
        return hipErrorInvalidValue;
    GemmArgs g{d.A, d.lda, d.B, d.ldb, d.M, d.N, d.K, d.k_split, d.bias, d.zprev, d.ldz, d.zout, d.ldzo,
               d.yout, d.ldy, d.ytout, d.ldyt, d.n_valid, d.lo_off};
    const bool small = d.N <= 32 || d.M <= 4096;
#define CP3_GEMM(TM, TN, EPI)                                                                              \
    hipLaunchKernelGGL((gemm_nt<TM, TN, EPI>), dim3((d.M + 32 * TM - 1) / (32 * TM), (d.N + 32 * TN - 1) / (32 * TN), \
                       d.k_split), dim3(256), 0, s, g)
    switch (d.epilogue) {
        case kEpiBiasGelu: if (small) CP3_GEMM(1, 1, kEpiBiasGelu); else CP3_GEMM(2, 2, kEpiBiasGelu); break;
        case kEpiBias: CP3_GEMM(1, 1, kEpiBias); break;
        case kEpiGeluGrad: if (small) CP3_GEMM(1, 1, kEpiGeluGrad); else CP3_GEMM(2, 2, kEpiGeluGrad); break;
        case kEpiPartial: CP3_GEMM(1, 1, kEpiPartial); break;
        default: return hipErrorInvalidValue;
    }
#undef CP3_GEMM
    return hipGetLastError();
}

hipError_t launch_featurize(const float* x, int M, const MlpScaling& sc, __bf16* f, __bf16* ft, hipStream_t s) {
    if (M <= 0) return hipSuccess;
    hipLaunchKernelGGL(featurize, dim3(blocks_for((long long)M * kMlpInPad)), dim3(256), 0, s, x, M, sc, f, ft);
    return hipGetLastError();
}

hipError_t launch_output_head(const float* z, int M, const float* t, float* y, __bf16* dz, __bf16* dzt,
                              float* loss_part, hipStream_t s) {
    if (M <= 0) return hipSuccess;
    hipLaunchKernelGGL(output_head, dim3(blocks_for((long long)M * kMlpOutPad)), dim3(256), 0, s, z, M, t, y, dz, dzt,
                       loss_part);
    return hipGetLastError();
}

int output_head_blocks(int M) { return int(blocks_for((long long)M * kMlpOutPad)); }

hipError_t launch_row_sums(const __bf16* a, int rows, int M, float* out, hipStream_t s) {
    hipLaunchKernelGGL(row_sums, dim3(rows), dim3(256), 0, s, a, M, out);
    return hipGetLastError();
}

hipError_t launch_sum_parts(const float* p, int parts, int n, float* out, hipStream_t s) {
    hipLaunchKernelGGL(sum_parts, dim3(blocks_for(n)), dim3(256), 0, s, p, parts, n, out);
    return hipGetLastError();
}

hipError_t launch_sum_floats(const float* p, int n, float* out, hipStream_t s) {
    hipLaunchKernelGGL(sum_floats, dim3(1), dim3(256), 0, s, p, n, out);
    return hipGetLastError();
}

hipError_t launch_adam_layer(float* w, float* m1, float* m2, const float* g, int N, int K, int n_valid, int k_valid,
                             const AdamStep& a, __bf16* wb, int ld_wb, int dup, __bf16* wtb, hipStream_t s) {
    if (wb && (dup > ld_wb || ld_wb > 2 * K || dup < 1)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(adam_layer, dim3(blocks_for((long long)N * K)), dim3(256), 0, s, w, m1, m2, g, N, K, n_valid,
                       k_valid, a, wb, ld_wb, dup, wtb);
    return hipGetLastError();
}

hipError_t launch_adam_vec(float* w, float* m1, float* m2, const float* g, int n, int n_valid, const AdamStep& a,
                           hipStream_t s) {
    hipLaunchKernelGGL(adam_vec, dim3(blocks_for(n)), dim3(256), 0, s, w, m1, m2, g, n, n_valid, a);
    return hipGetLastError();
}

}  // namespace cp3
