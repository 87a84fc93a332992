// CPU oracle for the CUDAProb3++ propagation path -- TEST INFRASTRUCTURE ONLY.
//
// This file restates, in plain C++ (double precision, OpenMP over cosines), the
// reference's host-side set-up and its per-(cosZ, E) Barger propagation.  It is
// the checker that the HIP library is compared against; nothing in the shipped
// library links, loads or calls it.  Only tests/, the smoke check and the
// cpu_baseline leg of the benchmark may use it.
//
// Every arithmetic expression keeps the operand order of the reference so that,
// compiled like the reference (g++ -O2, no -march, hence no FMA contraction),
// the results agree bit for bit with the reference CpuPropagator<double>.
// Citations are file:line in the reference tree (fkallen/CUDAProb3).
//
// Pins (see tests/test_oracle.py):
//   * bit-exact agreement with the compiled reference (oracle/_ref) and with the
//     committed golden fixtures generated from it (tests/golden/);
//   * closed forms: vacuum 3-flavour formula for cos>=0, layer matrix equal to
//     U exp(-i 2.534 (L/E) H) U^dagger (scipy expm), eigenvalues equal to
//     numpy eigvalsh of H, unitarity, chord-length sum rule, maxlayers counting.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <omp.h>

namespace {

constexpr double kGF = 1.52588e-4;     // constants.hpp:28  tworttwoGf
constexpr double kKm2Cm = 1.0e5;       // constants.hpp:31
constexpr double kREarth = 6371.0;     // constants.hpp:34
constexpr double kREarthCm = kREarth * kKm2Cm;  // constants.hpp:37
constexpr double kYe = 0.5;            // constants.hpp:40  density_convert
constexpr double kLoE = 2.534;         // physics.hpp:442   LoEfac

struct Cx { double re, im; };

// libm calls of the per-pair physics.  Built with -DCP3_ORACLE_LIBM_JITTER (liboracle_jitter.so)
// each result is moved by one ulp up, one ulp down or left alone, chosen by a hash of the
// argument: a stand-in for another faithfully rounded libm (such as the GPU's).  How far
// that moves the probabilities is the reference algorithm's own sensitivity to libm
// rounding, which sets the floor of any FP64 parity bar.
#ifdef CP3_ORACLE_LIBM_JITTER
inline double libm_jitter(double x, double y, std::uint64_t salt) {
    std::uint64_t b;
    std::memcpy(&b, &x, 8);
    b ^= salt;
    b ^= b >> 33;
    b *= 0xff51afd7ed558ccdULL;
    b ^= b >> 33;
    const int r = int(b % 3);
    return r == 0 ? y : std::nextafter(y, r == 1 ? 1e300 : -1e300);
}
inline double phys_sin(double x) { return libm_jitter(x, std::sin(x), 0x9e3779b97f4a7c15ULL); }
inline double phys_cos(double x) { return libm_jitter(x, std::cos(x), 0xc2b2ae3d27d4eb4fULL); }
inline double phys_acos(double x) { return libm_jitter(x, std::acos(x), 0x165667b19e3779f9ULL); }
#else
inline double phys_sin(double x) { return std::sin(x); }
inline double phys_cos(double x) { return std::cos(x); }
inline double phys_acos(double x) { return std::acos(x); }
#endif

// Parameter block shared by the physics routines (the reference keeps these in
// process globals, physics.hpp:121-124; here they travel by reference).
struct Params {
    Cx U[9];        // U(i,j) = U[i*3+j]            propagator.hpp:309-311
    double DM[9];   // DM(i,j) = DM[i*3+j]           propagator.hpp:313-315
    double ax[324]; // AXFAC(a,b,c,d,e)              physics.hpp:94/101
    int order[3];   // ORDER(i)                      physics.hpp:95/102
};

inline double& AX(Params& p, int a, int b, int c, int d, int e) {
    return p.ax[a * 108 + b * 36 + c * 12 + d * 4 + e];
}
inline double AXc(const Params& p, int a, int b, int c, int d, int e) {
    return p.ax[a * 108 + b * 36 + c * 12 + d * 4 + e];
}

// propagator.hpp:191-221  (PDG parametrisation of the PMNS matrix)
void mns_matrix(double t12, double t13, double t23, double dcp, Cx U[9]) {
    const double s12 = std::sin(t12), s13 = std::sin(t13), s23 = std::sin(t23);
    const double c12 = std::cos(t12), c13 = std::cos(t13), c23 = std::cos(t23);
    const double sd = std::sin(dcp), cd = std::cos(dcp);
    U[0] = {c12 * c13, 0.0};
    U[1] = {s12 * c13, 0.0};
    U[2] = {s13 * cd, -s13 * sd};
    U[3] = {-s12 * c23 - c12 * s23 * s13 * cd, -c12 * s23 * s13 * sd};
    U[4] = {c12 * c23 - s12 * s23 * s13 * cd, -s12 * s23 * s13 * sd};
    U[5] = {s23 * c13, 0.0};
    U[6] = {s12 * s23 - c12 * c23 * s13 * cd, -c12 * c23 * s13 * sd};
    U[7] = {-c12 * s23 - s12 * c23 * s13 * cd, -s12 * c23 * s13 * sd};
    U[8] = {c23 * c13, 0.0};
}

// propagator.hpp:226-248  (mass-squared differences with the 5e-9 degeneracy nudge)
void mass_matrix(double dm12sq, double dm23sq, double DM[9]) {
    double mv[3] = {0.0, dm12sq, dm12sq + dm23sq};
    const double nudge = 5.0e-9;
    if (dm12sq == 0.0) mv[0] -= nudge;
    if (dm23sq == 0.0) mv[2] += nudge;
    DM[0] = 0.0; DM[4] = 0.0; DM[8] = 0.0;
    DM[1] = mv[0] - mv[1];
    DM[3] = -DM[1];
    DM[2] = mv[0] - mv[2];
    DM[6] = -DM[2];
    DM[5] = mv[1] - mv[2];
    DM[7] = -DM[5];
}

// physics.hpp:157-174  (AXFAC: products U_ni * conj(U_mj) split in re/im parts)
void build_axfac(Params& p) {
    for (int n = 0; n < 3; n++)
        for (int m = 0; m < 3; m++)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    const Cx a = p.U[n * 3 + i], b = p.U[m * 3 + j];
                    AX(p, n, m, i, j, 0) = a.re * b.re + a.im * b.im;
                    AX(p, n, m, i, j, 1) = a.re * b.im - a.im * b.re;
                    AX(p, n, m, i, j, 2) = a.im * b.im + a.re * b.re;
                    AX(p, n, m, i, j, 3) = a.im * b.re - a.re * b.im;
                }
}

// physics.hpp:198-253  (vacuum cubic roots matched to DM(i,0) -> ORDER)
void vacuum_order(Params& p) {
    const double* D = p.DM;
    const double a = D[1] + D[2];
    const double b = D[1] * D[2];
    const double g = 0.0;
    double t = a * a - 3.0 * b;
    double arg = (2.0 * a * a * a - 9.0 * a * b + 27.0 * g) / (2.0 * std::sqrt(t * t * t));
    if (std::fabs(arg) > 1.0) arg = arg / std::fabs(arg);
    const double th0 = std::acos(arg) / 3.0;
    const double th1 = th0 - (2.0 * M_PI / 3.0);
    const double th2 = th0 + (2.0 * M_PI / 3.0);
    double root[3];
    root[0] = root[1] = root[2] = -(2.0 / 3.0) * std::sqrt(t);
    root[0] *= std::cos(th0); root[1] *= std::cos(th1); root[2] *= std::cos(th2);
    t = D[0] - a / 3.0;
    root[0] += t; root[1] += t; root[2] += t;
    for (int i = 0; i < 3; i++) {
        double best = std::fabs(D[i * 3] - root[0]);
        int k = 0;
        for (int j = 1; j < 3; j++) {
            const double d = std::fabs(D[i * 3] - root[j]);
            if (d < best) { k = j; best = d; }
        }
        p.order[i] = k;
    }
}

// physics.hpp:268-341  (matter eigenvalues, Barger et al. eqs. 21-22)
void matter_masses(const Params& p, double E, double rho, int antinu,
                   double mm[3][3], double mv[3][3]) {
    const double* D = p.DM;
    const Cx* U = p.U;
    const double fac = antinu ? kGF * E * rho : -kGF * E * rho;
    const double alpha = fac + D[1] + D[2];
    const double beta = D[1] * D[2] +
        fac * (D[1] * (1.0 - U[1].re * U[1].re - U[1].im * U[1].im) +
               D[2] * (1.0 - U[2].re * U[2].re - U[2].im * U[2].im));
    const double gamma = fac * D[1] * D[2] * (U[0].re * U[0].re + U[0].im * U[0].im);
    const double tmp = alpha * alpha - 3.0 * beta < 0 ? 0 : alpha * alpha - 3.0 * beta;
    const double argtmp = (2.0 * alpha * alpha * alpha - 9.0 * alpha * beta + 27.0 * gamma) /
                          (2.0 * std::sqrt(tmp * tmp * tmp));
    const double arg = std::fabs(argtmp) > 1.0 ? argtmp / std::fabs(argtmp) : argtmp;
    const double th0 = phys_acos(arg) / 3.0;
    const double th1 = th0 - (2.0 * M_PI / 3.0);
    const double th2 = th0 + (2.0 * M_PI / 3.0);
    double r[3];
    r[0] = -(2.0 / 3.0) * std::sqrt(tmp);
    r[1] = -(2.0 / 3.0) * std::sqrt(tmp);
    r[2] = -(2.0 / 3.0) * std::sqrt(tmp);
    r[0] *= phys_cos(th0);
    r[1] *= phys_cos(th1);
    r[2] *= phys_cos(th2);
    const double shift = D[0] - alpha / 3.0;
    r[0] += shift; r[1] += shift; r[2] += shift;
    double m[3];
    for (int i = 0; i < 3; i++) m[i] = r[p.order[i]];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            mm[i][j] = m[i] - m[j];
            mv[i][j] = m[i] - D[j * 3];
        }
}

// physics.hpp:346-431  (numerators of eq. 11: products of (2EH - M_j) over j != k)
void lagrange_products(const Params& p, double E, double rho, int antinu,
                       const double mv[3][3], const double mm[3][3], Cx prod[3][3][3]) {
    const Cx* U = p.U;
    const double fac = antinu ? kGF * E * rho : -kGF * E * rho;
    Cx h[3][3][3];
    for (int n = 0; n < 3; n++)
        for (int m = 0; m < 3; m++) {
            const double re = -fac * (U[n].re * U[m].re + U[n].im * U[m].im);
            const double im = -fac * (U[n].re * U[m].im - U[n].im * U[m].re);
            for (int j = 0; j < 3; j++) h[n][m][j] = {re, im};
        }
    for (int j = 0; j < 3; j++) {
        h[0][0][j].re -= mv[j][0];
        h[1][1][j].re -= mv[j][1];
        h[2][2][j].re -= mv[j][2];
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) prod[i][j][k] = {0.0, 0.0};
    // (k; a, b) = (0; 1, 2), (1; 2, 0), (2; 0, 1)
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            for (int k = 0; k < 3; k++) {
                prod[i][j][0].re += h[i][k][1].re * h[k][j][2].re - h[i][k][1].im * h[k][j][2].im;
                prod[i][j][0].im += h[i][k][1].re * h[k][j][2].im + h[i][k][1].im * h[k][j][2].re;
                prod[i][j][1].re += h[i][k][2].re * h[k][j][0].re - h[i][k][2].im * h[k][j][0].im;
                prod[i][j][1].im += h[i][k][2].re * h[k][j][0].im + h[i][k][2].im * h[k][j][0].re;
                prod[i][j][2].re += h[i][k][0].re * h[k][j][1].re - h[i][k][0].im * h[k][j][1].im;
                prod[i][j][2].im += h[i][k][0].re * h[k][j][1].im + h[i][k][0].im * h[k][j][1].re;
            }
            prod[i][j][0].re /= (mm[0][1] * mm[0][2]);
            prod[i][j][0].im /= (mm[0][1] * mm[0][2]);
            prod[i][j][1].re /= (mm[1][2] * mm[1][0]);
            prod[i][j][1].im /= (mm[1][2] * mm[1][0]);
            prod[i][j][2].re /= (mm[2][0] * mm[2][1]);
            prod[i][j][2].im /= (mm[2][0] * mm[2][1]);
        }
}

// physics.hpp:434-517  (X = sum_k exp(i phi_k) prod_k; A = U X U^dagger via AXFAC)
void layer_amplitude(const Params& p, double L, double E, double rho, int antinu,
                     const double mv[3][3], const double mm[3][3], Cx A[3][3]) {
    Cx prod[3][3][3];
    lagrange_products(p, E, rho, antinu, mv, mm, prod);
    Cx X[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) X[i][j] = {0.0, 0.0};
    for (int k = 0; k < 3; k++) {
        const double arg = -kLoE * mv[k][0] * L / E;
        const double s = phys_sin(arg);
        const double c = phys_cos(arg);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                X[i][j].re += c * prod[i][j][k].re - s * prod[i][j][k].im;
                X[i][j].im += c * prod[i][j][k].im + s * prod[i][j][k].re;
            }
    }
    for (int n = 0; n < 3; n++)
        for (int m = 0; m < 3; m++) A[n][m] = {0.0, 0.0};
    for (int n = 0; n < 3; n++)
        for (int m = 0; m < 3; m++)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    A[n][m].re += AXc(p, n, m, i, j, 0) * X[i][j].re + AXc(p, n, m, i, j, 1) * X[i][j].im;
                    A[n][m].im += AXc(p, n, m, i, j, 2) * X[i][j].im + AXc(p, n, m, i, j, 3) * X[i][j].re;
                }
}

// physics.hpp:522-531
void transition_matrix(const Params& p, int antinu, double E, double rho, double Lkm, Cx A[3][3]) {
    double mv[3][3], mm[3][3];
    matter_masses(p, E, rho, antinu, mm, mv);
    layer_amplitude(p, Lkm, E, rho, antinu, mv, mm, A);
}

// physics.hpp:536-548  (layer 0 is the vacuum atmosphere)
double layer_density(const double* rhos, int layer, int max_layer) {
    if (layer == 0) return 0.0;
    const int i = layer <= max_layer ? layer - 1 : 2 * max_layer - layer - 1;
    return rhos[i];
}

// physics.hpp:553-580  (chord length inside a shell, in cm)
double layer_distance(const double* radii, int layer, int max_layer, double path,
                      double earth_path, double cz) {
    if (cz >= 0) return path;
    if (layer == 0) return path - earth_path;
    const int i = layer >= max_layer ? -layer - 1 + 2 * max_layer : layer - 1;
    const double here = 2.0 * std::sqrt(radii[i] * radii[i] - (kREarth) * (kREarth) * (1 - cz * cz));
    if (i < max_layer - 1) {
        // the reference evaluates this for every layer; it is only used here, and reading
        // radii[i + 1] for the innermost layer would step past a model without an r = 0 row
        const double next = 2.0 * std::sqrt(radii[i + 1] * radii[i + 1] - (kREarth) * (kREarth) * (1 - cz * cz));
        return 0.5 * (here - next) * (kKm2Cm);
    }
    return here * (kKm2Cm);
}

// math.hpp:50-64  (C += A B)
void mat_mul_acc(const Cx A[3][3], const Cx B[3][3], Cx C[3][3]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                C[i][j].re += A[i][k].re * B[k][j].re - A[i][k].im * B[k][j].im;
                C[i][j].im += A[i][k].im * B[k][j].re + A[i][k].re * B[k][j].im;
            }
}

void set_zero(Cx M[3][3]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) M[i][j] = {0.0, 0.0};
}

void set_copy(const Cx A[3][3], Cx B[3][3]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) B[i][j] = A[i][j];
}

// physics.hpp:613-699, one (cosine, energy) pair; out[t], t = in*3+out_flavour
void propagate_pair(const Params& p, int antinu, double cz, double E, const double* radii,
                    const double* rhos, int max_layer, double h_cm, double out[9]) {
    const double path = std::sqrt((kREarthCm + h_cm) * (kREarthCm + h_cm) -
                                  (kREarthCm * kREarthCm) * (1 - cz * cz)) - kREarthCm * cz;
    const double earth_path = -2.0 * cz * kREarthCm;
    Cx T[3][3], mantle[3][3], fin[3][3], tmp[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) mantle[i][j] = {i == j ? 1.0 : 0.0, 0.0};
    for (int i = 0; i <= max_layer; i++) {
        const double dist = layer_distance(radii, i, max_layer, path, earth_path, cz);
        const double dens = layer_density(rhos, i, max_layer);
        transition_matrix(p, antinu, E, dens * kYe, dist / kKm2Cm, T);
        if (i == 0) {
            set_copy(T, fin);
        } else if (i < max_layer) {
            set_zero(tmp); mat_mul_acc(T, fin, tmp); set_copy(tmp, fin);
            set_zero(tmp); mat_mul_acc(mantle, T, tmp); set_copy(tmp, mantle);
        } else {
            set_zero(tmp); mat_mul_acc(T, fin, tmp); set_copy(tmp, fin);
        }
    }
    set_zero(tmp); mat_mul_acc(mantle, fin, tmp); set_copy(tmp, fin);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            const double re = fin[b][a].re, im = fin[b][a].im;
            out[a * 3 + b] = re * re + im * im;
        }
}

void load_params(Params& p, const double* U18, const double* DM9) {
    for (int i = 0; i < 9; i++) p.U[i] = {U18[2 * i], U18[2 * i + 1]};
    std::memcpy(p.DM, DM9, sizeof(double) * 9);
    build_axfac(p);
    vacuum_order(p);
}

}  // namespace

extern "C" {

// ---- host set-up, propagator.hpp ------------------------------------------------

void orc_mns_matrix(double t12, double t13, double t23, double dcp, double* U18) {
    Cx U[9];
    mns_matrix(t12, t13, t23, dcp, U);
    for (int i = 0; i < 9; i++) { U18[2 * i] = U[i].re; U18[2 * i + 1] = U[i].im; }
}

void orc_mass_matrix(double dm12sq, double dm23sq, double* DM9) { mass_matrix(dm12sq, dm23sq, DM9); }

// propagator.hpp:117-161.  Returns 0, or 1 (size mismatch), 2 (empty), 3 (radii order).
// Outputs are in descending-radius order.
int orc_set_density(const double* radii_in, const double* rhos_in, int n, double* radii_out,
                    double* rhos_out, double* coslimit) {
    if (n <= 0) return 2;
    bool flip = false;
    if (n >= 2) {
        const int sign = (radii_in[1] - radii_in[0] > 0 ? 1 : -1);
        for (int i = 1; i < n; i++)
            if ((radii_in[i] - radii_in[i - 1]) * sign < 0) return 3;
        flip = sign == 1;
    }
    for (int i = 0; i < n; i++) {
        const int s = flip ? n - 1 - i : i;
        radii_out[i] = radii_in[s];
        rhos_out[i] = rhos_in[s];
    }
    for (int i = 0; i < n; i++) {
        double x = -1 * std::sqrt(1 - (radii_out[i] * radii_out[i] / (kREarth * kREarth)));
        if (i == 0) x = 0;
        coslimit[i] = x;
    }
    return 0;
}

// propagator.hpp:301-307
void orc_maxlayers(const double* cosines, int n_c, const double* coslimit, int n_rows, int* maxlayers) {
    for (int ic = 0; ic < n_c; ic++) {
        int count = 0;
        for (int i = 0; i < n_rows; i++)
            if (cosines[ic] < coslimit[i]) count++;
        maxlayers[ic] = count;
    }
}

// physics.hpp:198-253
void orc_vacuum_order(const double* DM9, int* order3) {
    Params p;
    std::memcpy(p.DM, DM9, sizeof(double) * 9);
    vacuum_order(p);
    for (int i = 0; i < 3; i++) order3[i] = p.order[i];
}

// physics.hpp:130-152
void orc_axfac(const double* U18, double* ax324) {
    Params p;
    for (int i = 0; i < 9; i++) p.U[i] = {U18[2 * i], U18[2 * i + 1]};
    build_axfac(p);
    std::memcpy(ax324, p.ax, sizeof(double) * 324);
}

// ---- physics pieces ---------------------------------------------------------------

// physics.hpp:268-341: dmMatMat and dmMatVac (row-major 3x3 each); rho already x Ye.
void orc_matter_masses(const double* U18, const double* DM9, double E, double rho, int antinu,
                       double* mm9, double* mv9) {
    Params p;
    load_params(p, U18, DM9);
    double mm[3][3], mv[3][3];
    matter_masses(p, E, rho, antinu, mm, mv);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) { mm9[i * 3 + j] = mm[i][j]; mv9[i * 3 + j] = mv[i][j]; }
}

// physics.hpp:522-531: one layer's flavour-basis amplitude A[n][m] (re,im interleaved).
void orc_transition_matrix(const double* U18, const double* DM9, int antinu, double E, double rho,
                           double Lkm, double* A18) {
    Params p;
    load_params(p, U18, DM9);
    Cx A[3][3];
    transition_matrix(p, antinu, E, rho, Lkm, A);
    for (int n = 0; n < 3; n++)
        for (int m = 0; m < 3; m++) { A18[2 * (n * 3 + m)] = A[n][m].re; A18[2 * (n * 3 + m) + 1] = A[n][m].im; }
}

// physics.hpp:553-580 and 615-618: distances (cm) of layers 0..max_layer for one cosine.
void orc_layer_distances(const double* radii_desc, int max_layer, double cz, double h_cm, double* dist) {
    const double path = std::sqrt((kREarthCm + h_cm) * (kREarthCm + h_cm) -
                                  (kREarthCm * kREarthCm) * (1 - cz * cz)) - kREarthCm * cz;
    const double earth_path = -2.0 * cz * kREarthCm;
    for (int i = 0; i <= max_layer; i++) dist[i] = layer_distance(radii_desc, i, max_layer, path, earth_path, cz);
}

// physics.hpp:583-702 on the host (AoS result[(ic*n_e+ie)*9 + t]); inputs already set up.
void orc_calculate(int antinu, const double* cosines, int n_c, const double* energies, int n_e,
                   const double* radii_desc, const double* rhos_desc, const int* maxlayers,
                   double h_cm, const double* U18, const double* DM9, int threads, double* result) {
    Params p;
    load_params(p, U18, DM9);
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic)
    for (int ic = 0; ic < n_c; ic++) {
        for (int ie = 0; ie < n_e; ie++) {
            double* out = result + (static_cast<std::uint64_t>(ic) * n_e + ie) * 9;
            propagate_pair(p, antinu, cosines[ic], energies[ie], radii_desc, rhos_desc, maxlayers[ic], h_cm, out);
        }
    }
}

// The CpuPropagator<double> call sequence end to end (cpupropagator.hpp:83-95 after the
// setters of propagator.hpp): radii/rhos as in a density file (either order).
// Returns the orc_set_density status.
int orc_propagate(int antinu, const double* cosines, int n_c, const double* energies, int n_e,
                  const double* radii, const double* rhos, int n_rows, double h_km,
                  double t12, double t13, double t23, double dcp, double dm12sq, double dm23sq,
                  int threads, double* result) {
    double U18[18], DM9[9];
    orc_mns_matrix(t12, t13, t23, dcp, U18);
    orc_mass_matrix(dm12sq, dm23sq, DM9);
    double* rd = new double[n_rows];
    double* rh = new double[n_rows];
    double* cl = new double[n_rows];
    int* ml = new int[n_c > 0 ? n_c : 1];
    const int st = orc_set_density(radii, rhos, n_rows, rd, rh, cl);
    if (st == 0) {
        orc_maxlayers(cosines, n_c, cl, n_rows, ml);
        const double h_cm = h_km * 100000.0;
        orc_calculate(antinu, cosines, n_c, energies, n_e, rd, rh, ml, h_cm, U18, DM9, threads, result);
    }
    delete[] rd; delete[] rh; delete[] cl; delete[] ml;
    return st;
}

}  // extern "C"
