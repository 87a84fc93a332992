// FP64 pieces of the propagation path, compiled with -ffp-contract=off so that every
// product and sum rounds exactly like the reference's x86-64 host build (no FMA).
//
//  * propagate_plain_f64  -- one lane per (cosZ, E) pair, the full Barger step for every
//                            crossed layer (the reference's calculateKernel,
//                            physics.hpp:583-721), restructured for CDNA: one wavefront per
//                            cosine row x 64 energies, the AXFAC table in LDS, parameters
//                            by value in the kernel-argument segment.
//  * layer_geometry       -- per-cosine layer lengths (physics.hpp:553-580, 615-618).
//  * shell_spectrum<T>    -- per-(shell, energy) eigenvalues and flavour-basis projectors.
//                            A layer's amplitude is U (sum_k e^{i phi_k} P_k) U^dagger with
//                            P_k the reference's eq. 11 products (physics.hpp:346-431); the
//                            P_k depend on (E, rho) only, so they are computed once per
//                            (shell, energy) instead of once per (pair, layer).
#include "common.hpp"

namespace cp3 {
namespace {

struct Cd { double re, im; };

__device__ inline Cd Uof(const OscParams& p, int i, int j) {
    return Cd{p.U[2 * (3 * i + j)], p.U[2 * (3 * i + j) + 1]};
}

// Barger et al. eqs. 21-22: ordered matter eigenvalues (physics.hpp:268-341).
// mm[i][j] = m_i - m_j, mv[i][j] = m_i - DM(j,0).
__device__ inline void matter_masses(const OscParams& p, double E, double rho, double mm[3][3],
                                     double mv[3][3]) {
    const double* D = p.DM;
    const Cd u0 = Uof(p, 0, 0), u1 = Uof(p, 0, 1), u2 = Uof(p, 0, 2);
    const double fac = p.antinu ? kGF * E * rho : -kGF * E * rho;
    const double alpha = fac + D[1] + D[2];
    const double beta = D[1] * D[2] +
        fac * (D[1] * (1.0 - u1.re * u1.re - u1.im * u1.im) + D[2] * (1.0 - u2.re * u2.re - u2.im * u2.im));
    const double gamma = fac * D[1] * D[2] * (u0.re * u0.re + u0.im * u0.im);
    const double disc = alpha * alpha - 3.0 * beta;
    const double tmp = disc < 0 ? 0.0 : disc;
    const double argtmp = (2.0 * alpha * alpha * alpha - 9.0 * alpha * beta + 27.0 * gamma) /
                          (2.0 * sqrt(tmp * tmp * tmp));
    const double arg = fabs(argtmp) > 1.0 ? argtmp / fabs(argtmp) : argtmp;
    const double th0 = acos(arg) / 3.0;
    const double th[3] = {th0, th0 - (2.0 * kPi / 3.0), th0 + (2.0 * kPi / 3.0)};
    const double scale = -(2.0 / 3.0) * sqrt(tmp);
    const double shift = D[0] - alpha / 3.0;
    double root[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double r = scale;
        r *= cos(th[k]);
        root[k] = r + shift;
    }
    double m[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int o = p.order[i];
        m[i] = o == 0 ? root[0] : (o == 1 ? root[1] : root[2]);
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            mm[i][j] = m[i] - m[j];
            mv[i][j] = m[i] - D[3 * j];
        }
}

// Entry (n,m) of 2EH - M_j (physics.hpp:361-379): a Hermitian matrix whose off-diagonal
// part does not depend on j.
struct TwoEH {
    Cd base[3][3];
    double mv[3][3];
    __device__ inline Cd at(int n, int m, int j) const {
        Cd v = base[n][m];
        if (n == m) v.re -= mv[j][n];
        return v;
    }
};

__device__ inline void two_eh(const OscParams& p, double E, double rho, const double mv[3][3], TwoEH& h) {
    const double fac = p.antinu ? kGF * E * rho : -kGF * E * rho;
#pragma unroll
    for (int n = 0; n < 3; n++)
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const Cd a = Uof(p, 0, n), b = Uof(p, 0, m);
            h.base[n][m] = Cd{-fac * (a.re * b.re + a.im * b.im), -fac * (a.re * b.im - a.im * b.re)};
        }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) h.mv[i][j] = mv[i][j];
}

// Element (i,j) of the three eq. 11 projectors, (k; a, b) = (0; 1, 2), (1; 2, 0), (2; 0, 1)
// (physics.hpp:395-428), same summation order and divisions as the reference.
__device__ inline void projector_elem(const TwoEH& h, const double mm[3][3], int i, int j, Cd P[3]) {
    Cd acc[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int ab[3][2] = {{1, 2}, {2, 0}, {0, 1}};
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const Cd a = h.at(i, k, ab[q][0]);
            const Cd b = h.at(k, j, ab[q][1]);
            acc[q].re += a.re * b.re - a.im * b.im;
            acc[q].im += a.re * b.im + a.im * b.re;
        }
    }
    const double d0 = mm[0][1] * mm[0][2];
    const double d1 = mm[1][2] * mm[1][0];
    const double d2 = mm[2][0] * mm[2][1];
    P[0] = Cd{acc[0].re / d0, acc[0].im / d0};
    P[1] = Cd{acc[1].re / d1, acc[1].im / d1};
    P[2] = Cd{acc[2].re / d2, acc[2].im / d2};
}

// AXFAC(n,m,i,j,0..3) (physics.hpp:135-146): U_ni conj(U_mj) split in parts.
__device__ inline void axfac(const OscParams& p, int n, int m, int i, int j, double f[4]) {
    const Cd a = Uof(p, n, i), b = Uof(p, m, j);
    f[0] = a.re * b.re + a.im * b.im;
    f[1] = a.re * b.im - a.im * b.re;
    f[2] = a.im * b.im + a.re * b.re;
    f[3] = a.im * b.re - a.re * b.im;
}

__device__ inline double layer_density(const double* rhos, int layer, int max_layer) {
    if (layer == 0) return 0.0;
    const int i = layer <= max_layer ? layer - 1 : 2 * max_layer - layer - 1;
    return rhos[i];
}

// physics.hpp:553-580, result in cm.
__device__ inline double layer_distance(const double* radii, int layer, int max_layer, double path,
                                        double earth_path, double cz) {
    if (cz >= 0) return path;
    if (layer == 0) return path - earth_path;
    const int i = layer >= max_layer ? -layer - 1 + 2 * max_layer : layer - 1;
    const double chord = (kREarth) * (kREarth) * (1 - cz * cz);
    const double here = 2.0 * sqrt(radii[i] * radii[i] - chord);
    if (i < max_layer - 1) {
        const double next = 2.0 * sqrt(radii[i + 1] * radii[i + 1] - chord);
        return 0.5 * (here - next) * (kKm2Cm);
    }
    return here * (kKm2Cm);
}

__device__ inline double path_length(double cz, double h_cm) {
    return sqrt((kREarthCm + h_cm) * (kREarthCm + h_cm) - (kREarthCm * kREarthCm) * (1 - cz * cz)) -
           kREarthCm * cz;
}

// C = A B, with the summation order of math.hpp:50-64.
__device__ inline void mat_mul(const Cd A[3][3], const Cd B[3][3], Cd C[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            Cd c{0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                c.re += A[i][k].re * B[k][j].re - A[i][k].im * B[k][j].im;
                c.im += A[i][k].im * B[k][j].re + A[i][k].re * B[k][j].im;
            }
            C[i][j] = c;
        }
}

// One constant-density layer's flavour-basis amplitude (physics.hpp:434-531) with the
// AXFAC table read from LDS.
__device__ inline void layer_amplitude(const OscParams& p, const double* __restrict__ ax, double E,
                                       double rho, double Lkm, Cd A[3][3]) {
    double mm[3][3], mv[3][3];
    matter_masses(p, E, rho, mm, mv);
    TwoEH h;
    two_eh(p, E, rho, mv, h);
    double c[3], s[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double arg = -kLoE * mv[k][0] * Lkm / E;
        s[k] = sin(arg);
        c[k] = cos(arg);
    }
    Cd X[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            Cd P[3];
            projector_elem(h, mm, i, j, P);
            Cd x{0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                x.re += c[k] * P[k].re - s[k] * P[k].im;
                x.im += c[k] * P[k].im + s[k] * P[k].re;
            }
            X[i][j] = x;
        }
#pragma unroll
    for (int n = 0; n < 3; n++)
#pragma unroll
        for (int m = 0; m < 3; m++) {
            Cd a{0.0, 0.0};
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const double* f = ax + (n * 108 + m * 36 + i * 12 + j * 4);
                    a.re += f[0] * X[i][j].re + f[1] * X[i][j].im;
                    a.im += f[2] * X[i][j].im + f[3] * X[i][j].re;
                }
            A[n][m] = a;
        }
}

__global__ __launch_bounds__(kBlock) void propagate_plain_f64(GridDesc g, OscParams p, double* __restrict__ out) {
    __shared__ double ax[324];
    for (int t = threadIdx.x; t < 81; t += blockDim.x) {
        const int n = t / 27, m = (t / 9) % 3, i = (t / 3) % 3, j = t % 3;
        axfac(p, n, m, i, j, ax + 4 * t);
    }
    __syncthreads();

    const unsigned lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int tiles = (g.n_e + kWave - 1) / kWave;
    const int row_groups = (g.n_c + 3) / 4;
    const int tile = lb / row_groups;
    const int ic = (lb % row_groups) * 4 + threadIdx.x / kWave;
    if (tile >= tiles || ic >= g.n_c) return;
    const int ie = tile * kWave + threadIdx.x % kWave;
    if (ie >= g.n_e) return;

    const double cz = g.cosines[ic];
    const double E = g.energies[ie];
    const int L = g.maxlayers[ic];
    const double path = path_length(cz, g.h_cm);
    const double earth_path = -2.0 * cz * kREarthCm;

    Cd T[3][3], fin[3][3], mantle[3][3], tmp[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) mantle[i][j] = Cd{i == j ? 1.0 : 0.0, 0.0};

    for (int i = 0; i <= L; i++) {
        const double dist = layer_distance(g.radii, i, L, path, earth_path, cz);
        const double dens = layer_density(g.rhos, i, L);
        layer_amplitude(p, ax, E, dens * kYe, dist / kKm2Cm, T);
        if (i == 0) {
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) fin[a][b] = T[a][b];
        } else {
            mat_mul(T, fin, tmp);
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) fin[a][b] = tmp[a][b];
            if (i < L) {
                mat_mul(mantle, T, tmp);
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int b = 0; b < 3; b++) mantle[a][b] = tmp[a][b];
            }
        }
    }
    mat_mul(mantle, fin, tmp);

    const std::size_t plane = std::size_t(g.n_c) * g.n_e;
    double* o = out + std::size_t(ic) * g.n_e + ie;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const Cd v = tmp[b][a];
            o[(a * 3 + b) * plane] = v.re * v.re + v.im * v.im;
        }
}

__global__ __launch_bounds__(kBlock) void layer_geometry(GridDesc g, double* __restrict__ geom) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int slots = g.rows + 1;
    if (t >= g.n_c * slots) return;
    const int ic = t / slots, i = t % slots;
    const int L = g.maxlayers[ic];
    double v = 0.0;
    if (i <= L) {
        const double cz = g.cosines[ic];
        const double path = path_length(cz, g.h_cm);
        const double earth_path = -2.0 * cz * kREarthCm;
        v = layer_distance(g.radii, i, L, path, earth_path, cz) / kKm2Cm;
    }
    geom[t] = v;
}

// Per (shell s, energy e): Q_k = U P_k U^dagger (Hermitian, packed to 9 reals per k) and the
// phase coefficient of eigenvalue k.  kRevolutions: R = -2.534 lambda_k / (2 pi E), the
// per-pair phase in turns is R * L; otherwise R = -2.534 lambda_k and the phase is R*L/E.
template <typename T, bool kRevolutions>
__global__ __launch_bounds__(kBlock) void shell_spectrum(GridDesc g, OscParams p, int ne_pad,
                                                         T* __restrict__ Q, double* __restrict__ R) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (g.rows + 1) * ne_pad) return;
    const int s = t / ne_pad, e = t % ne_pad;
    const double E = g.energies[e < g.n_e ? e : g.n_e - 1];
    const double rho = (s == 0 ? 0.0 : g.rhos[s - 1]) * kYe;

    double mm[3][3], mv[3][3];
    matter_masses(p, E, rho, mm, mv);
    TwoEH h;
    two_eh(p, E, rho, mv, h);
    Cd P[3][3][3];  // [k][i][j]
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            Cd v[3];
            projector_elem(h, mm, i, j, v);
            P[0][i][j] = v[0]; P[1][i][j] = v[1]; P[2][i][j] = v[2];
        }
    const std::size_t st = ne_pad;
    T* q = Q + std::size_t(s) * 27 * st + e;
    double* r = R + std::size_t(s) * 3 * st + e;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        // Q_k[n][m] for (n,m) in {00, 11, 22, 01, 02, 12}
        const int nm[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
        double packed[9];
#pragma unroll
        for (int w = 0; w < 6; w++) {
            const int n = nm[w][0], m = nm[w][1];
            Cd a{0.0, 0.0};
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    double f[4];
                    axfac(p, n, m, i, j, f);
                    a.re += f[0] * P[k][i][j].re + f[1] * P[k][i][j].im;
                    a.im += f[2] * P[k][i][j].im + f[3] * P[k][i][j].re;
                }
            if (w < 3) {
                packed[w] = a.re;
            } else {
                packed[3 + 2 * (w - 3)] = a.re;
                packed[4 + 2 * (w - 3)] = a.im;
            }
        }
#pragma unroll
        for (int f = 0; f < 9; f++) q[(k * 9 + f) * st] = T(packed[f]);
        const double c = -kLoE * mv[k][0];
        r[k * st] = kRevolutions ? c / (E * (2.0 * kPi)) : c;
    }
}

}  // namespace

hipError_t launch_plain_f64(const GridDesc& g, const OscParams& p, double* out, hipStream_t s) {
    if (g.n_c == 0 || g.n_e == 0) return hipSuccess;
    const unsigned tiles = (g.n_e + kWave - 1) / kWave;
    const unsigned blocks = round_up(tiles * ((g.n_c + 3) / 4), kXcds);
    hipLaunchKernelGGL(propagate_plain_f64, dim3(blocks), dim3(kBlock), 0, s, g, p, out);
    return hipGetLastError();
}

hipError_t launch_geometry(const GridDesc& g, double* geom, hipStream_t s) {
    const unsigned n = unsigned(g.n_c) * (g.rows + 1);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(layer_geometry, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, g, geom);
    return hipGetLastError();
}

hipError_t launch_spectrum_f32(const GridDesc& g, const OscParams& p, int ne_pad, float* Q, double* R,
                               hipStream_t s) {
    const unsigned n = unsigned(g.rows + 1) * ne_pad;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL((shell_spectrum<float, true>), dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, g, p,
                       ne_pad, Q, R);
    return hipGetLastError();
}

hipError_t launch_spectrum_f64(const GridDesc& g, const OscParams& p, int ne_pad, double* Q, double* R,
                               hipStream_t s) {
    const unsigned n = unsigned(g.rows + 1) * ne_pad;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL((shell_spectrum<double, false>), dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, g,
                       p, ne_pad, Q, R);
    return hipGetLastError();
}

}  // namespace cp3
