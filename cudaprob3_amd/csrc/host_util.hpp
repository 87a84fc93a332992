// Host-side helpers shared by the C ABI translation units.
#pragma once

#include "../../include/cp3hip.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

namespace cp3host {

// Message of the last failing call on this thread (cp3h_last_error).
inline thread_local std::string g_error;

inline int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

#define CP3H_HIP(call)                                                                           \
    do {                                                                                         \
        hipError_t err_ = (call);                                                                \
        if (err_ != hipSuccess)                                                                  \
            return ::cp3host::fail(CP3H_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(err_)); \
    } while (0)

// Vacuum cubic roots matched to DM(i,0) (physics.hpp:198-253), in the caller's precision:
// order[i] is the matter root that continues vacuum mass i.
template <typename T>
void vacuum_order(const T* DM, int order[3]) {
    const T a = DM[1] + DM[2];
    const T b = DM[1] * DM[2];
    const T g = 0.0;
    T t = a * a - T(3.0) * b;
    T arg = (T(2.0) * a * a * a - T(9.0) * a * b + T(27.0) * g) / (T(2.0) * std::sqrt(t * t * t));
    if (std::fabs(arg) > T(1.0)) arg = arg / std::fabs(arg);
    const T th0 = std::acos(arg) / T(3.0);
    const T th[3] = {th0, T(th0 - (2.0 * M_PI / 3.0)), T(th0 + (2.0 * M_PI / 3.0))};
    T root[3];
    for (int k = 0; k < 3; k++) {
        root[k] = T(-(2.0 / 3.0) * std::sqrt(t));
        root[k] *= std::cos(th[k]);
    }
    t = DM[0] - a / T(3.0);
    for (int k = 0; k < 3; k++) root[k] += t;
    for (int i = 0; i < 3; i++) {
        T best = std::fabs(DM[i * 3] - root[0]);
        int k = 0;
        for (int j = 1; j < 3; j++) {
            const T d = std::fabs(DM[i * 3] - root[j]);
            if (d < best) { k = j; best = d; }
        }
        order[i] = k;
    }
}

// Real basis of the mixed-precision kernels.  The vacuum part of the flavour Hamiltonian,
// M = U diag(0, DM(1,0), DM(2,0)) U^dagger, and the matter term (a multiple of e_0 e_0^T) are
// brought to real symmetric form by one unitary W that fixes e_0: W = diag(1, W'), with W''s
// first column along (M_10, M_20) and its second column's phase chosen so that the (1,2) entry
// turns real.  Then every layer Hamiltonian is W H_r W^dagger with H_r = Hr - fac e_0 e_0^T
// real symmetric, and a layer amplitude is W S W^dagger with S complex symmetric (S^T = S).
// Works for any unitary U (no phase convention assumed).
//   Hr[6]  = (H00, H11, H22, H01, H02, H12) of W^dagger M W (H02 = 0 by construction)
//   Wmu[4] = row 1 of W: (Re W11, Im W11, Re W12, Im W12) (row 0 is e_0^T)
inline void real_basis(const double* U18, const double* DM9, double Hr[6], double Wmu[4]) {
    struct C { double re, im; };
    auto u = [&](int i, int j) { return C{U18[2 * (3 * i + j)], U18[2 * (3 * i + j) + 1]}; };
    C M[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            C acc{0.0, 0.0};
            for (int n = 1; n < 3; n++) {                 // U_in DM(n,0) conj(U_jn)
                const C a = u(i, n), b = u(j, n);
                acc.re += DM9[3 * n] * (a.re * b.re + a.im * b.im);
                acc.im += DM9[3 * n] * (a.im * b.re - a.re * b.im);
            }
            M[i][j] = acc;
        }
    // a = (M_10, M_20) / |.|, b orthogonal to a
    const double nx = std::sqrt(M[1][0].re * M[1][0].re + M[1][0].im * M[1][0].im +
                                M[2][0].re * M[2][0].re + M[2][0].im * M[2][0].im);
    C a[2] = {{1.0, 0.0}, {0.0, 0.0}};
    if (nx > 0.0) a[0] = C{M[1][0].re / nx, M[1][0].im / nx}, a[1] = C{M[2][0].re / nx, M[2][0].im / nx};
    C b[2] = {{-a[1].re, a[1].im}, {a[0].re, -a[0].im}};    // (-conj a1, conj a0)
    // quadratic forms of the 2x2 block M' = M[1..2][1..2]
    auto form = [&](const C* x, const C* y) {               // x^dagger M' y
        C acc{0.0, 0.0};
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) {
                const C m = M[1 + i][1 + j];
                const C my{m.re * y[j].re - m.im * y[j].im, m.re * y[j].im + m.im * y[j].re};
                acc.re += x[i].re * my.re + x[i].im * my.im;
                acc.im += x[i].re * my.im - x[i].im * my.re;
            }
        return acc;
    };
    const C m = form(a, b);
    const double nm = std::sqrt(m.re * m.re + m.im * m.im);
    if (nm > 0.0) {                                         // b <- b e^{-i arg m}
        const C ph{m.re / nm, -m.im / nm};
        for (auto& z : b) z = C{z.re * ph.re - z.im * ph.im, z.re * ph.im + z.im * ph.re};
    }
    Hr[0] = M[0][0].re;
    Hr[1] = form(a, a).re;
    Hr[2] = form(b, b).re;
    Hr[3] = nx;
    Hr[4] = 0.0;
    Hr[5] = nm;
    Wmu[0] = a[0].re; Wmu[1] = a[0].im;
    Wmu[2] = b[0].re; Wmu[3] = b[0].im;
}

// The single-grid mixed kernel's vacuum constants U_xk conj(U_yk) (k = 1, 2; x, y in (e, mu)),
// from the caller's U in FP64, rounded to float: c[(k - 1) * 8 + (x * 2 + y) * 2 + {0: re, 1: im}].
inline void vacuum_block(const double* U18, float c[16]) {
    for (int k = 1; k < 3; k++)
        for (int x = 0; x < 2; x++)
            for (int y = 0; y < 2; y++) {
                const double ar = U18[2 * (3 * x + k)], ai = U18[2 * (3 * x + k) + 1];
                const double br = U18[2 * (3 * y + k)], bi = U18[2 * (3 * y + k) + 1];
                const int i = (k - 1) * 8 + (x * 2 + y) * 2;
                c[i] = float(ar * br + ai * bi);
                c[i + 1] = float(ai * br - ar * bi);
            }
}

}  // namespace cp3host
