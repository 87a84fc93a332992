// FP64 pieces of the propagation path, compiled with -ffp-contract=off so that every
// product and sum rounds exactly like the reference's x86-64 host build (no FMA).
//
//  * propagate_plain_f64  -- one lane per (cosZ, E) pair, the full Barger step for every
//                            crossed layer (the reference's calculateKernel,
//                            physics.hpp:583-721), restructured for CDNA: one wavefront per
//                            cosine row x 64 energies, the AXFAC table in LDS, parameters
//                            by value in the kernel-argument segment.
//  * layer_geometry       -- per-cosine layer lengths (physics.hpp:553-580, 615-618).
//  * shell_spectrum_f64   -- per-(shell, energy) eigenvalues and flavour-basis projectors.
//                            A layer's amplitude is U (sum_k e^{i phi_k} P_k) U^dagger with
//                            P_k the reference's eq. 11 products (physics.hpp:346-431); the
//                            P_k depend on (E, rho) only, so they are computed once per
//                            (shell, energy) instead of once per (pair, layer).
//  * spectrum_cell_mixed  -- the same spectrum for the mixed kernels, as real eigenvectors in
//                            the real basis of OscParams::Hr (set-up kernels grid_setup_mixed
//                            and batch_setup).
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

// sin x and cos x for 0 <= x <= pi/3 in FP64: their Taylor polynomials to degrees 17 and 18 in
// Horner form (truncation below 2e-17 on that range; about one ulp in all).
__device__ inline void sincos_third(double x, double& sn, double& cs) {
    const double x2 = x * x;
    double ps = 2.8114572543455206e-15;
    ps = ps * x2 + -7.647163731819816e-13;
    ps = ps * x2 + 1.6059043836821613e-10;
    ps = ps * x2 + -2.505210838544172e-08;
    ps = ps * x2 + 2.7557319223985893e-06;
    ps = ps * x2 + -0.0001984126984126984;
    ps = ps * x2 + 0.008333333333333333;
    ps = ps * x2 + -0.16666666666666666;
    sn = x + x * (x2 * ps);
    double pc = -1.5619206968586225e-16;
    pc = pc * x2 + 4.779477332387385e-14;
    pc = pc * x2 + -1.1470745597729725e-11;
    pc = pc * x2 + 2.08767569878681e-09;
    pc = pc * x2 + -2.755731922398589e-07;
    pc = pc * x2 + 2.48015873015873e-05;
    pc = pc * x2 + -0.001388888888888889;
    pc = pc * x2 + 0.041666666666666664;
    pc = pc * x2 + -0.5;
    cs = 1.0 + x2 * pc;
}

// The same ordered matter eigenvalues for the mixed kernels' tables, with a shorter dependency
// chain (the set-up kernels are latency-bound): the three roots' cosines cos(th0 -+ 2 pi / 3) by
// angle addition from one sine and cosine of th0 in [0, pi/3] (sincos_third) instead of three
// library cosines, and no division where a reciprocal square root does.  Equal to
// matter_masses to a few ulps of the largest eigenvalue; not used by the FP64 kernels, whose
// bar follows the reference's own rounding.
__device__ inline void matter_masses_fast(const OscParams& p, double E, double rho, double mm[3][3],
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
    const double rs = rsqrt(tmp);                        // 1 / sqrt(tmp)
    const double argtmp = (2.0 * alpha * alpha * alpha - 9.0 * alpha * beta + 27.0 * gamma) * (0.5 * rs * rs * rs);
    const double arg = fmin(1.0, fmax(-1.0, argtmp));
    double s0, c0;
    sincos_third(acos(arg) / 3.0, s0, c0);
    const double scale = -(2.0 / 3.0) * sqrt(tmp);      // off the acos chain
    const double shift = D[0] - alpha / 3.0;
    const double half_c = -0.5 * c0, rt3_s = 0.8660254037844386 * s0;
    const double root[3] = {scale * c0 + shift, scale * (half_c + rt3_s) + shift, scale * (half_c - rt3_s) + shift};
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

// One constant-density layer's flavour-basis amplitude (physics.hpp:434-531) with the AXFAC
// table read from LDS.  Each X[i][j] = sum_k e^{i phi_k} P_k[i][j] is added into every A[n][m]
// as soon as it is formed, so X is never held whole; every A[n][m] still receives its terms
// in the reference's (i, j) order, which keeps its rounding.
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
#pragma unroll
    for (int n = 0; n < 3; n++)
#pragma unroll
        for (int m = 0; m < 3; m++) A[n][m] = Cd{0.0, 0.0};
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
#pragma unroll
            for (int n = 0; n < 3; n++)
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    const double* f = ax + (n * 108 + m * 36 + i * 12 + j * 4);
                    A[n][m].re += f[0] * x.re + f[1] * x.im;
                    A[n][m].im += f[2] * x.im + f[3] * x.re;
                }
        }
}

// The two running products of every lane of a workgroup, parked in LDS while a layer's
// amplitude is formed: element e of lane t at park[e][t] (fin in e = 0..17, mantle in 18..35,
// (row * 3 + column) * 2 + {re, im}).  Each lane touches only its own column, so no barrier is
// needed.  With them in registers the kernel needed 316 registers per lane (one wavefront per
// SIMD); parked, the amplitude's own working set fits 256 (two wavefronts per SIMD).
constexpr int kPlainBlock = 128;                 // 2 wavefronts per workgroup
constexpr int kPlainRows = kPlainBlock / kWave;  // cosine rows per workgroup
constexpr int kFin = 0, kMantle = 18;

struct Parked {
    double (*p)[kPlainBlock];
    int t;
    __device__ Cd get(int base, int a, int b) const {
        const int e = base + (a * 3 + b) * 2;
        return Cd{p[e][t], p[e + 1][t]};
    }
    __device__ void put(int base, int a, int b, Cd v) const {
        const int e = base + (a * 3 + b) * 2;
        p[e][t] = v.re;
        p[e + 1][t] = v.im;
    }
};

// B <- A B for the parked B, one column at a time (column j of the product needs column j of B
// only), with the summation order of math.hpp:50-64.
__device__ inline void mul_left(const Cd A[3][3], const Parked& P, int base) {
#pragma unroll
    for (int j = 0; j < 3; j++) {
        Cd b[3], col[3];
#pragma unroll
        for (int k = 0; k < 3; k++) b[k] = P.get(base, k, j);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            Cd c{0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                c.re += A[i][k].re * b[k].re - A[i][k].im * b[k].im;
                c.im += A[i][k].im * b[k].re + A[i][k].re * b[k].im;
            }
            col[i] = c;
        }
#pragma unroll
        for (int i = 0; i < 3; i++) P.put(base, i, j, col[i]);
    }
}

// A <- A B for the parked A, one row at a time, same summation order.
__device__ inline void mul_right(const Parked& P, int base, const Cd B[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
        Cd a[3], row[3];
#pragma unroll
        for (int k = 0; k < 3; k++) a[k] = P.get(base, i, k);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            Cd c{0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                c.re += a[k].re * B[k][j].re - a[k].im * B[k][j].im;
                c.im += a[k].im * B[k][j].re + a[k].re * B[k][j].im;
            }
            row[j] = c;
        }
#pragma unroll
        for (int j = 0; j < 3; j++) P.put(base, i, j, row[j]);
    }
}

__global__ __launch_bounds__(kPlainBlock) void propagate_plain_f64(GridDesc g, OscParams p, double* __restrict__ out) {
    __shared__ double ax[324];
    __shared__ double park[36][kPlainBlock];
    for (int t = threadIdx.x; t < 81; t += blockDim.x) {
        const int n = t / 27, m = (t / 9) % 3, i = (t / 3) % 3, j = t % 3;
        axfac(p, n, m, i, j, ax + 4 * t);
    }
    __syncthreads();

    const unsigned lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int tiles = (g.n_e + kWave - 1) / kWave;
    const int row_groups = (g.n_c + kPlainRows - 1) / kPlainRows;
    const int tile = lb / row_groups;
    const int ic = (lb % row_groups) * kPlainRows + threadIdx.x / kWave;
    if (tile >= tiles || ic >= g.n_c) return;
    const int ie = tile * kWave + threadIdx.x % kWave;
    if (ie >= g.n_e) return;

    const double cz = g.cosines[ic];
    const double E = g.energies[ie];
    const int L = crossed_shells(g.maxlayers, ic, g.rows);
    const double path = path_length(cz, g.h_cm);
    const double earth_path = -2.0 * cz * kREarthCm;

    // physics.hpp:645-677: fin = T_i fin for every layer, mantle = mantle T_i for 1 <= i < L
    const Parked P{park, int(threadIdx.x)};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) P.put(kMantle, i, j, Cd{i == j ? 1.0 : 0.0, 0.0});

    Cd T[3][3];
    for (int i = 0; i <= L; i++) {
        const double dist = layer_distance(g.radii, i, L, path, earth_path, cz);
        const double dens = layer_density(g.rhos, i, L);
        // the AXFAC table is loop-invariant: hoisted out of the layer loop, its 324 doubles
        // would stay live in registers across the loop and spill; an offset the compiler cannot
        // see through keeps its LDS reads inside each layer step
        int opaque_zero = 0;
        asm volatile("" : "+s"(opaque_zero));
        layer_amplitude(p, ax + opaque_zero, E, dens * kYe, dist / kKm2Cm, T);
        if (i == 0) {
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) P.put(kFin, a, b, T[a][b]);
        } else {
            mul_left(T, P, kFin);
            if (i < L) mul_right(P, kMantle, T);
        }
    }

    // total = mantle fin; P(a -> b) = |total[b][a]|^2, each element formed where it is stored
    const std::size_t plane = std::size_t(g.n_c) * g.n_e;
    double* o = out + std::size_t(ic) * g.n_e + ie;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        Cd f[3];
#pragma unroll
        for (int k = 0; k < 3; k++) f[k] = P.get(kFin, k, a);
#pragma unroll
        for (int b = 0; b < 3; b++) {
            Cd v{0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const Cd m = P.get(kMantle, b, k);
                v.re += m.re * f[k].re - m.im * f[k].im;
                v.im += m.im * f[k].re + m.re * f[k].im;
            }
            o[(a * 3 + b) * plane] = v.re * v.re + v.im * v.im;
        }
    }
}

// Length in km of layer i of cosine row ic (0 beyond the row's crossed layers).
__device__ inline double geometry_cell(const GridDesc& g, int ic, int i) {
    const int L = crossed_shells(g.maxlayers, ic, g.rows);
    if (i > L) return 0.0;
    const double cz = g.cosines[ic];
    const double path = path_length(cz, g.h_cm);
    const double earth_path = -2.0 * cz * kREarthCm;
    return layer_distance(g.radii, i, L, path, earth_path, cz) / kKm2Cm;
}

__global__ __launch_bounds__(kBlock) void layer_geometry(GridDesc g, double* __restrict__ geom) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int slots = g.rows + 1;
    if (t >= g.n_c * slots) return;
    geom[t] = geometry_cell(g, t / slots, t % slots);
}

// One flavour-basis projector of one (energy, density) layer Hamiltonian:
// Q_k = U P_k U^dagger (P_k the eq. 11 product of physics.hpp:395-428 with the (k; a, b)
// pairing (0; 1, 2), (1; 2, 0), (2; 0, 1)), Hermitian, packed as 9 reals:
// (Q00, Q11, Q22, Re Q01, Im Q01, Re Q02, Im Q02, Re Q12, Im Q12).  Each entry of P_k is summed
// in the reference's order (the same operations as projector_elem's entry k).  Also returns the
// phase coefficient rate = -2.534 lambda_k.
template <int k>
__device__ inline void flavour_projector(const OscParams& p, double E, double rho, double Q[9], double& rate) {
    double mm[3][3], mv[3][3];
    matter_masses(p, E, rho, mm, mv);
    TwoEH h;
    two_eh(p, E, rho, mv, h);
    constexpr int a = k == 0 ? 1 : (k == 1 ? 2 : 0), b = k == 0 ? 2 : (k == 1 ? 0 : 1);
    const double d = mm[k][a] * mm[k][b];
    Cd P[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            Cd acc{0.0, 0.0};
#pragma unroll
            for (int l = 0; l < 3; l++) {
                const Cd x = h.at(i, l, a);
                const Cd y = h.at(l, j, b);
                acc.re += x.re * y.re - x.im * y.im;
                acc.im += x.re * y.im + x.im * y.re;
            }
            P[i][j] = Cd{acc.re / d, acc.im / d};
        }
    const int nm[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
#pragma unroll
    for (int w = 0; w < 6; w++) {
        const int n = nm[w][0], m = nm[w][1];
        Cd acc{0.0, 0.0};
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                double f[4];
                axfac(p, n, m, i, j, f);
                acc.re += f[0] * P[i][j].re + f[1] * P[i][j].im;
                acc.im += f[2] * P[i][j].im + f[3] * P[i][j].re;
            }
        if (w < 3) {
            Q[w] = acc.re;
        } else {
            Q[3 + 2 * (w - 3)] = acc.re;
            Q[4 + 2 * (w - 3)] = acc.im;
        }
    }
    rate = -kLoE * mv[k][0];
}

__device__ inline void slot_inputs(const GridDesc& g, int s, int e, double& E, double& rho) {
    E = g.energies[e < g.n_e ? e : g.n_e - 1];
    rho = (s == 0 ? 0.0 : g.rhos[s - 1]) * kYe;
}

// FP64 table, per (slot s, energy e): all three Q_k (Q[(s*27 + k*9 + f)*ne_pad + e]) and the
// phase coefficients R[(s*3 + k)*ne_pad + e] = -2.534 lambda_k (the per-pair phase is R*L/E,
// the reference's expression).  Each cell is formed by three lanes of three wavefronts, one
// projector each (wavefront w takes projector w mod 3 of 64 consecutive cells, so the choice is
// wave-uniform; the cubic and 2EH - M are formed by all three): the kernel is latency-bound, and
// one lane forming all three projectors took three times the dependency chain and 256
// registers.  ne_pad is a multiple of 64, so the 64 cells of a wavefront share the slot.
__global__ __launch_bounds__(kBlock) void shell_spectrum_f64(GridDesc g, OscParams p, int ne_pad,
                                                             double* __restrict__ Q, double* __restrict__ R) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int w = t / kWave;
    const int k = __builtin_amdgcn_readfirstlane(w % 3);
    const int cell = (w / 3) * kWave + t % kWave;
    if (cell >= (g.rows + 1) * ne_pad) return;
    const int s = cell / ne_pad, e = cell % ne_pad;
    double E, rho;
    slot_inputs(g, s, e, E, rho);
    double q[9], rate;
    if (k == 0)
        flavour_projector<0>(p, E, rho, q, rate);
    else if (k == 1)
        flavour_projector<1>(p, E, rho, q, rate);
    else
        flavour_projector<2>(p, E, rho, q, rate);
    const std::size_t st = ne_pad;
#pragma unroll
    for (int f = 0; f < 9; f++) Q[(std::size_t(s) * 27 + k * 9 + f) * st + e] = q[f];
    R[(std::size_t(s) * 3 + k) * st + e] = rate;
}

// Mixed-precision table, per (slot s, energy e), in the real basis of OscParams::Hr / Wmu.
// There the layer Hamiltonian H = Hr - fac e_0 e_0^T is real symmetric (fac as in the
// reference, physics.hpp:277-282), so its eigenvectors v_k are real, and the layer amplitude up
// to its global phase is
//   S = I + (e^{i dphi_1} - 1) v_1 v_1^T + (e^{i dphi_2} - 1) v_2 v_2^T,
// dphi_k = -2.534 (lambda_k - lambda_0) L / E (sum_k v_k v_k^T = I).  The projectors are the
// reference's eq. 11 construction with its eigenvalues and (k; a, b) pairing
// (physics.hpp:395-428), Q_k = (H - lambda_a)(H - lambda_b) / ((lambda_k - lambda_a)(lambda_k -
// lambda_b)), here of a real 3x3 matrix; v_k is the column of Q_k through its largest diagonal
// entry, normalised (Q_k = v_k v_k^T).  Stored: v_1, v_2 in float as two float4 groups
// Qm[((s*2 + g)*ne_pad + e)*4 + c] (packed index g*4 + c: 0-2 v_1, 3-5 v_2, 6-7 zero), and
// dphi_k / L in half turns per km as Rm[(s*ne_pad + e)*2 + k - 1], double.
__device__ inline void real_eigenvectors_12(const OscParams& p, double E, double rho, double v[2][3],
                                            double mm[3][3], double mv[3][3]) {
    matter_masses_fast(p, E, rho, mm, mv);
    const double fac = p.antinu ? kGF * E * rho : -kGF * E * rho;
    double H[3][3];
    H[0][0] = p.Hr[0] - fac; H[1][1] = p.Hr[1]; H[2][2] = p.Hr[2];
    H[0][1] = H[1][0] = p.Hr[3];
    H[0][2] = H[2][0] = p.Hr[4];
    H[1][2] = H[2][1] = p.Hr[5];
    const int pair[2][2] = {{2, 0}, {0, 1}};    // k = 1: (a, b) = (2, 0); k = 2: (0, 1)
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
        const int k = kk + 1, a = pair[kk][0], b = pair[kk][1];
        const double la = mv[a][0], lb = mv[b][0];       // lambda - DM(0,0), DM(0,0) = 0
        const double inv = 1.0 / (mm[k][a] * mm[k][b]);
        double Q[3][3];
#pragma unroll
        for (int n = 0; n < 3; n++)
#pragma unroll
            for (int m = n; m < 3; m++) {
                double acc = 0.0;
#pragma unroll
                for (int l = 0; l < 3; l++) {
                    const double x = H[n][l] - (n == l ? la : 0.0);
                    const double y = H[l][m] - (l == m ? lb : 0.0);
                    acc += x * y;
                }
                Q[n][m] = Q[m][n] = acc * inv;
            }
        // the column through the largest diagonal entry, picked by selects: an index into Q that
        // is not a compile-time constant makes the compiler keep Q in LDS
        const bool j1 = Q[1][1] > Q[0][0];
        const double d01 = j1 ? Q[1][1] : Q[0][0];
        const bool j2 = Q[2][2] > d01;
        const double djj = j2 ? Q[2][2] : d01;
        const double s = djj > 0.0 ? rsqrt(djj) : 0.0;
#pragma unroll
        for (int n = 0; n < 3; n++) v[kk][n] = (j2 ? Q[n][2] : (j1 ? Q[n][1] : Q[n][0])) * s;
    }
}

// Real orthonormal eigenbasis O = [v_1 x v_2, v_1, v_2] (det +1) from v_1, v_2, as columns
// o[k][component].
__device__ inline void eigenbasis(const double v[2][3], double o[3][3]) {
    o[0][0] = v[0][1] * v[1][2] - v[0][2] * v[1][1];
    o[0][1] = v[0][2] * v[1][0] - v[0][0] * v[1][2];
    o[0][2] = v[0][0] * v[1][1] - v[0][1] * v[1][0];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        o[1][c] = v[0][c];
        o[2][c] = v[1][c];
    }
}

// Spectrum cell (slot s, energy e).  Without kBasis one lane decomposes slot s and stores v_1,
// v_2 and the phase rates (half = 0).  With kBasis the cell has two adjacent lanes of one
// wavefront (half = the lane's parity): lane 0 does the above; lane 1 decomposes slot s - 1
// (O_0 = I for s = 1) and, with lane 0's eigenbasis, stores the basis change M_s = O_s^T O_{s-1}
// (row r: (o_r . o'_c)_c).  Both lanes reach the exchange; each does one decomposition, so the
// cell costs the latency of one.  M_0 is not used.
template <bool kBasis>
__device__ inline void spectrum_cell_mixed(const GridDesc& g, const OscParams& p, int ne_pad, int s, int e, int half,
                                           float* __restrict__ Qm, float* __restrict__ Mm, double* __restrict__ Rm) {
    const std::size_t st = ne_pad;
    double o[3][3];
    if (half == 0 || s >= 2) {
        double E, rho;
        slot_inputs(g, s - half, e, E, rho);
        double v[2][3], mm[3][3], mv[3][3];
        real_eigenvectors_12(p, E, rho, v, mm, mv);
        eigenbasis(v, o);
        if (half == 0) {
            float4* dst = reinterpret_cast<float4*>(Qm);
            dst[(std::size_t(s) * kMixedQGroups + 0) * st + e] =
                make_float4(float(v[0][0]), float(v[0][1]), float(v[0][2]), float(v[1][0]));
            dst[(std::size_t(s) * kMixedQGroups + 1) * st + e] =
                make_float4(float(v[1][1]), float(v[1][2]), 0.0f, 0.0f);
            const double half_turn = 1.0 / (E * kPi);         // radians -> half turns
            Rm[(std::size_t(s) * st + e) * 2 + 0] = -kLoE * mm[1][0] * half_turn;
            Rm[(std::size_t(s) * st + e) * 2 + 1] = -kLoE * mm[2][0] * half_turn;
        }
    } else {
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) o[r][c] = r == c ? 1.0 : 0.0;
    }
    if constexpr (!kBasis) return;
    double cur[3][3];                 // lane 0's eigenbasis O_s, on both lanes
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) cur[r][c] = __shfl_xor(o[r][c], 1);
    if (half == 1) {
        float4* mdst = reinterpret_cast<float4*>(Mm);
#pragma unroll
        for (int r = 0; r < 3; r++) {
            float m[3];
#pragma unroll
            for (int c = 0; c < 3; c++)
                m[c] = float(cur[r][0] * o[c][0] + cur[r][1] * o[c][1] + cur[r][2] * o[c][2]);
            mdst[(std::size_t(s) * kMixedMGroups + r) * st + e] = make_float4(m[0], m[1], m[2], 0.0f);
        }
    }
}

// Set-up of one mixed-precision grid in one launch: cells [0, G) are layer lengths
// (cosine, layer), [G, G + S) spectrum cells (slot, energy).
__global__ __launch_bounds__(kBlock) void grid_setup_mixed(GridDesc g, OscParams p, int ne_pad, double* __restrict__ geom,
                                                           float* __restrict__ Qm, float* __restrict__ Mm,
                                                           double* __restrict__ Rm) {
    const int slots = g.rows + 1;
    const int G = g.n_c * slots, S = slots * ne_pad;
    const int G2 = (G + 1) & ~1;       // spectrum lanes start even: a cell's two lanes share a wavefront
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < G) {
        geom[t] = geometry_cell(g, t / slots, t % slots);
    } else if (t >= G2 && t < G2 + 2 * S) {
        const int cell = (t - G2) >> 1;
        spectrum_cell_mixed<true>(g, p, ne_pad, cell / ne_pad, cell % ne_pad, (t - G2) & 1, Qm, Mm, Rm);
    }
}

// Set index of work cell t given the per-set prefix sums off[0..n] (off[0] = 0).
__device__ inline int find_set(const long long* __restrict__ off, int n, long long t) {
    int lo = 0, hi = n;          // off[lo] <= t < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// Batched set-up of many parameter sets in one launch: work cells [0, G) are layer lengths
// (set, cosine, layer), [G, G + S) spectrum cells (set, slot, energy) without basis changes (the
// batched pair kernel forms them from the eigenvectors).  Also resets the pair kernel's work queue.
__global__ __launch_bounds__(kBlock) void batch_setup(BatchDesc b) {
    const long long G = b.geom_prefix[b.n_sets], S = b.spec_prefix[b.n_sets];
    if (blockIdx.x == 0 && threadIdx.x < kXcds) b.next_item[threadIdx.x * kQueueStride] = 0;   // pair kernel follows
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < G + S;
         t += (long long)gridDim.x * blockDim.x) {
        if (t < G) {
            const int k = find_set(b.geom_prefix, b.n_sets, t);
            const BatchSet& set = b.sets[k];
            const GridDesc g = batch_grid(b, set);
            const long long c = t - b.geom_prefix[k];
            const int slots = set.rows + 1;
            b.geom[set.geom_off + c] = geometry_cell(g, int(c / slots), int(c % slots));
        } else {
            const long long u = t - G;
            const int k = find_set(b.spec_prefix, b.n_sets, u);
            const BatchSet& set = b.sets[k];
            const GridDesc g = batch_grid(b, set);
            const long long c = u - b.spec_prefix[k];
            spectrum_cell_mixed<false>(g, set.p, set.ne_pad, int(c / set.ne_pad), int(c % set.ne_pad), 0,
                                       reinterpret_cast<float*>(b.Q + set.tab_off * kMixedQGroups), nullptr,
                                       reinterpret_cast<double*>(b.R + set.tab_off));
        }
    }
}

}  // namespace

hipError_t launch_plain_f64(const GridDesc& g, const OscParams& p, double* out, hipStream_t s) {
    if (g.n_c == 0 || g.n_e == 0) return hipSuccess;
    const unsigned tiles = (g.n_e + kWave - 1) / kWave;
    const unsigned blocks = round_up(tiles * ((g.n_c + kPlainRows - 1) / kPlainRows), kXcds);
    hipLaunchKernelGGL(propagate_plain_f64, dim3(blocks), dim3(kPlainBlock), 0, s, g, p, out);
    return hipGetLastError();
}

hipError_t launch_geometry(const GridDesc& g, double* geom, hipStream_t s) {
    const unsigned n = unsigned(g.n_c) * (g.rows + 1);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(layer_geometry, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, g, geom);
    return hipGetLastError();
}

hipError_t launch_setup_mixed(const GridDesc& g, const OscParams& p, int ne_pad, double* geom, float* Qm, float* Mm,
                              double* Rm, hipStream_t s) {
    if (g.n_c == 0 || g.n_e == 0) return hipSuccess;
    const unsigned G = unsigned(g.n_c) * (g.rows + 1);
    const unsigned n = ((G + 1) & ~1u) + 2u * unsigned(g.rows + 1) * ne_pad;
    hipLaunchKernelGGL(grid_setup_mixed, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, g, p, ne_pad, geom,
                       Qm, Mm, Rm);
    return hipGetLastError();
}

hipError_t launch_batch_setup(const BatchDesc& b, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(batch_setup, dim3(blocks), dim3(kBlock), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_spectrum_f64(const GridDesc& g, const OscParams& p, int ne_pad, double* Q, double* R,
                               hipStream_t s) {
    const unsigned n = 3u * unsigned(g.rows + 1) * ne_pad;
    if (n == 0 || g.n_e == 0) return hipSuccess;
    hipLaunchKernelGGL(shell_spectrum_f64, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, g, p, ne_pad, Q,
                       R);
    return hipGetLastError();
}

}  // namespace cp3
