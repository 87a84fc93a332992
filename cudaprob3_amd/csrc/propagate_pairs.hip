// Table-driven per-pair propagation.
//
// A layer's amplitude matrix is A = sum_k e^{i phi_k} Q_k with Q_k = U P_k U^dagger the
// flavour-basis projectors of the layer Hamiltonian (the reference's eq. 11 products,
// physics.hpp:346-517, rearranged by linearity).  Q_k depends on (energy, shell) only and
// comes from the spectrum table; per pair and layer only the three phases
// phi_k = -2.534 lambda_k L / E, the sum over k, and the running products
// (physics.hpp:645-677) remain.
//
// Mixed precision (pairs_mixed): the phase is formed and reduced modulo one turn in FP64
// (it reaches ~10^3 rad on core-crossing baselines at 0.05 GeV, where FP32 would lose
// 1e-4 rad), then the hardware sin/cos (v_sin_f32/v_cos_f32, argument in turns) and all
// matrix arithmetic run in FP32.  FP64 (pairs_f64): the same structure in double with the
// reference's phase expression ((-2.534 lambda_k) L) / E.
//
// Mapping: one wavefront = one cosine row x 64 consecutive energies (lane = energy); four
// wavefronts of a workgroup take four consecutive rows of the same energy tile, and
// consecutive logical workgroups stay on one XCD so that an energy tile's table slice is
// reused from that XCD's L2.  Layer lengths of the row are wave-uniform scalar loads.
#include "common.hpp"

namespace cp3 {
namespace {

template <typename T> struct Cx { T re, im; };

template <typename T>
__device__ inline void mul3(const Cx<T> A[3][3], const Cx<T> B[3][3], Cx<T> C[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            T re = A[i][0].re * B[0][j].re - A[i][0].im * B[0][j].im;
            T im = A[i][0].im * B[0][j].re + A[i][0].re * B[0][j].im;
#pragma unroll
            for (int k = 1; k < 3; k++) {
                re += A[i][k].re * B[k][j].re - A[i][k].im * B[k][j].im;
                im += A[i][k].im * B[k][j].re + A[i][k].re * B[k][j].im;
            }
            C[i][j] = Cx<T>{re, im};
        }
}

// A = sum_k (c_k + i s_k) Q_k for Hermitian Q_k packed as (Q00, Q11, Q22, Q01, Q02, Q12).
template <typename T>
__device__ inline void assemble(const T q[27], const T c[3], const T s[3], Cx<T> A[3][3]) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
        T re = c[0] * q[d], im = s[0] * q[d];
#pragma unroll
        for (int k = 1; k < 3; k++) {
            re += c[k] * q[k * 9 + d];
            im += s[k] * q[k * 9 + d];
        }
        A[d][d] = Cx<T>{re, im};
    }
    const int idx[3][2] = {{0, 1}, {0, 2}, {1, 2}};
#pragma unroll
    for (int w = 0; w < 3; w++) {
        T ure = 0, uim = 0, lre = 0, lim = 0;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const T a = q[k * 9 + 3 + 2 * w], b = q[k * 9 + 4 + 2 * w];
            const T ca = c[k] * a, sb = s[k] * b, cb = c[k] * b, sa = s[k] * a;
            ure += ca - sb;
            uim += cb + sa;
            lre += ca + sb;
            lim += sa - cb;
        }
        A[idx[w][0]][idx[w][1]] = Cx<T>{ure, uim};
        A[idx[w][1]][idx[w][0]] = Cx<T>{lre, lim};
    }
}

struct PairArgs {
    const int* maxlayers;
    const double* energies;
    const double* geom;      // [n_c][rows + 1] km
    const void* Q;
    const double* R;
    void* out;               // SoA [9][n_c][n_e]
    int n_c, n_e, rows, ne_pad;
};

// kMixed: T = float, R in turns; otherwise T = double, R = -2.534 lambda.
template <typename T, bool kMixed>
__global__ __launch_bounds__(kBlock) void pairs_kernel(PairArgs a) {
    const unsigned lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int tiles = (a.n_e + kWave - 1) / kWave;
    const int row_groups = (a.n_c + 3) / 4;
    const int tile = lb / row_groups;
    const int ic = (lb % row_groups) * 4 + threadIdx.x / kWave;
    if (tile >= tiles || ic >= a.n_c) return;
    const int ie = tile * kWave + threadIdx.x % kWave;   // < ne_pad: table reads stay in bounds

    const int L = a.maxlayers[ic];
    const double* gl = a.geom + std::size_t(ic) * (a.rows + 1);
    const T* Q = static_cast<const T*>(a.Q);
    const std::size_t st = a.ne_pad;
    double E = 0.0;
    if (!kMixed) E = a.energies[ie < a.n_e ? ie : a.n_e - 1];

    Cx<T> fin[3][3], mantle[3][3], A[3][3], tmp[3][3];
    for (int i = 0; i <= L; i++) {
        const double dist = gl[i];
        const T* q = Q + std::size_t(i) * 27 * st + ie;
        const double* r = a.R + std::size_t(i) * 3 * st + ie;
        T qv[27];
#pragma unroll
        for (int f = 0; f < 27; f++) qv[f] = q[f * st];
        T c[3], s[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (kMixed) {
                double turns = r[k * st] * dist;
                turns -= rint(turns);
                const float x = float(turns);
                s[k] = __builtin_amdgcn_sinf(x);
                c[k] = __builtin_amdgcn_cosf(x);
            } else {
                const double arg = r[k * st] * dist / E;
                double sv, cv;
                sincos(arg, &sv, &cv);
                s[k] = T(sv);
                c[k] = T(cv);
            }
        }
        assemble<T>(qv, c, s, A);
        if (i == 0) {
#pragma unroll
            for (int x = 0; x < 3; x++)
#pragma unroll
                for (int y = 0; y < 3; y++) fin[x][y] = A[x][y];
        } else {
            mul3<T>(A, fin, tmp);
#pragma unroll
            for (int x = 0; x < 3; x++)
#pragma unroll
                for (int y = 0; y < 3; y++) fin[x][y] = tmp[x][y];
            if (i < L) {
                if (i == 1) {
#pragma unroll
                    for (int x = 0; x < 3; x++)
#pragma unroll
                        for (int y = 0; y < 3; y++) mantle[x][y] = A[x][y];
                } else {
                    mul3<T>(mantle, A, tmp);
#pragma unroll
                    for (int x = 0; x < 3; x++)
#pragma unroll
                        for (int y = 0; y < 3; y++) mantle[x][y] = tmp[x][y];
                }
            }
        }
    }
    if (L >= 2) {
        mul3<T>(mantle, fin, tmp);
#pragma unroll
        for (int x = 0; x < 3; x++)
#pragma unroll
            for (int y = 0; y < 3; y++) fin[x][y] = tmp[x][y];
    }
    if (ie >= a.n_e) return;
    const std::size_t plane = std::size_t(a.n_c) * a.n_e;
    T* o = static_cast<T*>(a.out) + std::size_t(ic) * a.n_e + ie;
#pragma unroll
    for (int x = 0; x < 3; x++)
#pragma unroll
        for (int y = 0; y < 3; y++) {
            const Cx<T> v = fin[y][x];
            o[(x * 3 + y) * plane] = v.re * v.re + v.im * v.im;
        }
}

template <typename T, bool kMixed>
hipError_t launch_pairs(const GridDesc& g, int ne_pad, const double* geom, const void* Q, const double* R,
                        void* out, hipStream_t s) {
    if (g.n_c == 0 || g.n_e == 0) return hipSuccess;
    PairArgs a{g.maxlayers, g.energies, geom, Q, R, out, g.n_c, g.n_e, g.rows, ne_pad};
    const unsigned tiles = (g.n_e + kWave - 1) / kWave;
    const unsigned blocks = round_up(tiles * ((g.n_c + 3) / 4), kXcds);
    hipLaunchKernelGGL((pairs_kernel<T, kMixed>), dim3(blocks), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_pairs_mixed(const GridDesc& g, int ne_pad, const double* geom, const float* Q,
                              const double* R, float* out, hipStream_t s) {
    return launch_pairs<float, true>(g, ne_pad, geom, Q, R, out, s);
}

hipError_t launch_pairs_f64(const GridDesc& g, int ne_pad, const double* geom, const double* Q,
                            const double* R, double* out, hipStream_t s) {
    return launch_pairs<double, false>(g, ne_pad, geom, Q, R, out, s);
}

}  // namespace cp3
