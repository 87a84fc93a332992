// The (cosZ, E) pairs of the mixed-precision propagation, shared by the single-grid and the
// batched kernels: propagate_rows_mixed<kRows> takes kRows cosine rows with the same number of
// crossed shells at one energy per lane.  Layer algebra and table layout: see
// propagate_pairs.hip.
//   L    : shells crossed (maxlayers of the rows)
//   gl[k]: layer lengths of row k in km, gl[k][0..L]
//   Q, R : the spectrum table (float4 groups / double2 phase rates in quarter turns per km),
//          row stride st
//   ie   : energy index (< ne_pad)
// Writes p[k][in][out] = P(in -> out) of row k.
#pragma once
#include "common.hpp"

namespace cp3 {

template <typename T> struct Cx { T re, im; };

__device__ inline float clamp01(float x) { return fminf(1.0f, fmaxf(0.0f, x)); }

// acc += a * b for complex a, b as four fused multiply-adds (written out so the compiler
// cannot split them into separate multiplies and adds).
__device__ inline void cmac(Cx<float>& acc, Cx<float> a, Cx<float> b) {
    acc.re = __builtin_fmaf(a.re, b.re, acc.re);
    acc.re = __builtin_fmaf(-a.im, b.im, acc.re);
    acc.im = __builtin_fmaf(a.re, b.im, acc.im);
    acc.im = __builtin_fmaf(a.im, b.re, acc.im);
}

__device__ inline Cx<float> cmul(Cx<float> a, Cx<float> b) {
    Cx<float> r{a.re * b.re, a.re * b.im};
    r.re = __builtin_fmaf(-a.im, b.im, r.re);
    r.im = __builtin_fmaf(a.im, b.re, r.im);
    return r;
}

// e^{i phi} - 1 for a phase given in quarter turns (phi = t * pi/2), as (cos phi - 1, sin phi).
// Cody-Waite style reduction in FP64: q = rint(t), r = t - q in [-1/2, 1/2] is exact, so the
// phase reaches FP32 only as the small remainder z = r * pi/2 in [-pi/4, pi/4]; there
// sin z and cos z - 1 are Taylor polynomials (next terms below 3e-9 relative) evaluated in
// FP32 by Horner's rule, and the quadrant q mod 4 rotates the result.
__device__ inline void phase_factor_m1(double t, float& cm1, float& sn) {
    // every fused multiply-add is written out and nothing else may be fused, so that each
    // inlined copy (one row, two rows, batched) rounds identically
#pragma clang fp contract(off)
    const double q = rint(t);
    const float z = float((t - q) * 1.5707963267948966);
    const float z2 = z * z;
    float ps = __builtin_fmaf(z2, 2.75573192e-6f, -1.98412698e-4f);
    ps = __builtin_fmaf(z2, ps, 8.33333333e-3f);
    ps = __builtin_fmaf(z2, ps, -1.66666667e-1f);
    const float s0 = __builtin_fmaf(z * z2, ps, z);
    float pc = __builtin_fmaf(z2, -2.75573192e-7f, 2.48015873e-5f);
    pc = __builtin_fmaf(z2, pc, -1.38888889e-3f);
    pc = __builtin_fmaf(z2, pc, 4.16666667e-2f);
    pc = __builtin_fmaf(z2, pc, -0.5f);
    const float c0m1 = z2 * pc;
    switch (int(q) & 3) {
        case 0: cm1 = c0m1; sn = s0; break;                  // cos z, sin z
        case 1: cm1 = -s0 - 1.0f; sn = c0m1 + 1.0f; break;   // -sin z, cos z
        case 2: cm1 = -c0m1 - 2.0f; sn = -s0; break;         // -cos z, -sin z
        default: cm1 = s0 - 1.0f; sn = -(c0m1 + 1.0f); break; // sin z, -cos z
    }
}

// One shell's table entries at this lane's energy: Q_1, Q_2 (packed, see propagate_pairs.hip)
// and the two phase rates.  They depend on (shell, energy) only, so rows that cross the same
// shells share them.
struct LayerTable {
    float q[20];
    double2 r;
};

__device__ inline __attribute__((always_inline)) void load_table(const float4* __restrict__ qp,
                                                                const double2* __restrict__ rp, std::size_t st,
                                                                int i, LayerTable& t) {
#pragma unroll
    for (int grp = 0; grp < 5; grp++) {
        const float4 v = qp[(std::size_t(i) * 5 + grp) * st];
        t.q[4 * grp] = v.x; t.q[4 * grp + 1] = v.y; t.q[4 * grp + 2] = v.z; t.q[4 * grp + 3] = v.w;
    }
    t.r = rp[std::size_t(i) * st];
}

// The layer amplitude up to its global phase, A = I + w_1 Q_1 + w_2 Q_2 with
// w_k = e^{i dphi_k} - 1; Q Hermitian, so A_ji uses conj(Q_ij).
__device__ inline __attribute__((always_inline)) void layer_amplitude_mixed(const LayerTable& t, double dist,
                                                                           Cx<float> A[3][3]) {
    const float* q = t.q;
    float wr[2], wi[2];
    phase_factor_m1(t.r.x * dist, wr[0], wi[0]);
    phase_factor_m1(t.r.y * dist, wr[1], wi[1]);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        A[d][d].re = __builtin_fmaf(wr[1], q[9 + d], __builtin_fmaf(wr[0], q[d], 1.0f));
        A[d][d].im = __builtin_fmaf(wi[1], q[9 + d], wi[0] * q[d]);
    }
    const int idx[3][2] = {{0, 1}, {0, 2}, {1, 2}};
#pragma unroll
    for (int w = 0; w < 3; w++) {
        const float a0 = q[3 + 2 * w], b0 = q[4 + 2 * w], a1 = q[12 + 2 * w], b1 = q[13 + 2 * w];
        // w_k (a + i b) and w_k (a - i b), summed over k
        float ure = wr[0] * a0, uim = wr[0] * b0, lre = ure, lim = wi[0] * a0;
        ure = __builtin_fmaf(-wi[0], b0, ure);
        uim = __builtin_fmaf(wi[0], a0, uim);
        lre = __builtin_fmaf(wi[0], b0, lre);
        lim = __builtin_fmaf(-wr[0], b0, lim);
        ure = __builtin_fmaf(wr[1], a1, ure);
        ure = __builtin_fmaf(-wi[1], b1, ure);
        uim = __builtin_fmaf(wr[1], b1, uim);
        uim = __builtin_fmaf(wi[1], a1, uim);
        lre = __builtin_fmaf(wr[1], a1, lre);
        lre = __builtin_fmaf(wi[1], b1, lre);
        lim = __builtin_fmaf(wi[1], a1, lim);
        lim = __builtin_fmaf(-wr[1], b1, lim);
        A[idx[w][0]][idx[w][1]] = Cx<float>{ure, uim};
        A[idx[w][1]][idx[w][0]] = Cx<float>{lre, lim};
    }
}

// F <- A F (F: columns 0-1 of the forward product)
__device__ inline __attribute__((always_inline)) void left_mul_cols(const Cx<float> A[3][3], Cx<float> F[3][2]) {
    Cx<float> G[3][2];
#pragma unroll
    for (int x = 0; x < 3; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) {
            Cx<float> acc = cmul(A[x][0], F[0][y]);
            cmac(acc, A[x][1], F[1][y]);
            cmac(acc, A[x][2], F[2][y]);
            G[x][y] = acc;
        }
#pragma unroll
    for (int x = 0; x < 3; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) F[x][y] = G[x][y];
}

// M <- M A (M: rows 0-1 of the outgoing mantle product)
__device__ inline __attribute__((always_inline)) void right_mul_rows(Cx<float> M[2][3], const Cx<float> A[3][3]) {
    Cx<float> N[2][3];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 3; y++) {
            Cx<float> acc = cmul(M[x][0], A[0][y]);
            cmac(acc, M[x][1], A[1][y]);
            cmac(acc, M[x][2], A[2][y]);
            N[x][y] = acc;
        }
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 3; y++) M[x][y] = N[x][y];
}

// The path's layers in the reference's order (physics.hpp:645-677): layer 0 (atmosphere), then
// shells 1..L-1 on the way in, which enter both the forward product and the mantle product
// of the way out, then the innermost shell L, crossed once; total = mantle x forward.  The
// first two layers and the last are peeled off the loop, so its body has no branches; L is
// wave-uniform, so every branch here is a scalar one.
//
// kRows cosine rows with the same L are propagated together at this lane's energy: they cross
// the same shells, so each shell's table entries are loaded once for all of them (only the
// layer lengths, hence the phases, differ), which halves the table reads per pair for kRows = 2.
template <int kRows>
__device__ inline void propagate_rows_mixed(int L, const double* const* gl, const float4* __restrict__ Q,
                                            const double2* __restrict__ R, std::size_t st, int ie,
                                            float (*p)[3][3]) {
    const float4* qp = Q + ie;
    const double2* rp = R + ie;
    LayerTable t;
    Cx<float> A[kRows][3][3];
    Cx<float> F[kRows][3][2];   // columns 0-1 of T_i ... T_1 T_atm
    Cx<float> M[kRows][2][3];   // rows 0-1 of T_1 ... T_i
    load_table(qp, rp, st, 0, t);
#pragma unroll
    for (int k = 0; k < kRows; k++) {
        layer_amplitude_mixed(t, gl[k][0], A[k]);
#pragma unroll
        for (int x = 0; x < 3; x++)
#pragma unroll
            for (int y = 0; y < 2; y++) F[k][x][y] = A[k][x][y];
    }
    if (L >= 2) {
        load_table(qp, rp, st, 1, t);
#pragma unroll
        for (int k = 0; k < kRows; k++) {
            layer_amplitude_mixed(t, gl[k][1], A[k]);
            left_mul_cols(A[k], F[k]);
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 3; y++) M[k][x][y] = A[k][x][y];
        }
        // one row: unrolled twice, so that the running products alternate registers instead of
        // being copied; two rows carry enough independent work as they are
        auto shell = [&](int i) {
            load_table(qp, rp, st, i, t);
#pragma unroll
            for (int k = 0; k < kRows; k++) {
                layer_amplitude_mixed(t, gl[k][i], A[k]);
                left_mul_cols(A[k], F[k]);
                right_mul_rows(M[k], A[k]);
            }
        };
        if (kRows == 1) {
#pragma unroll 2
            for (int i = 2; i < L; i++) shell(i);
        } else {
#pragma unroll 1
            for (int i = 2; i < L; i++) shell(i);
        }
    }
    if (L >= 1) {
        load_table(qp, rp, st, L, t);
#pragma unroll
        for (int k = 0; k < kRows; k++) {
            layer_amplitude_mixed(t, gl[k][L], A[k]);
            left_mul_cols(A[k], F[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < kRows; k++) {
        // upper-left 2x2 block S[out][in] of the total amplitude
        Cx<float> S[2][2];
        if (L >= 2) {
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++) {
                    Cx<float> acc = cmul(M[k][x][0], F[k][0][y]);
                    cmac(acc, M[k][x][1], F[k][1][y]);
                    cmac(acc, M[k][x][2], F[k][2][y]);
                    S[x][y] = acc;
                }
        } else {
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++) S[x][y] = F[k][x][y];
        }
        // p[in][out] = |S[out][in]|^2; the third row and column from unitarity
        float (&pk)[3][3] = p[k];
#pragma unroll
        for (int in = 0; in < 2; in++)
#pragma unroll
            for (int o = 0; o < 2; o++)
                pk[in][o] = __builtin_fmaf(S[o][in].re, S[o][in].re, S[o][in].im * S[o][in].im);
        pk[0][2] = clamp01(1.0f - pk[0][0] - pk[0][1]);
        pk[1][2] = clamp01(1.0f - pk[1][0] - pk[1][1]);
        pk[2][0] = clamp01(1.0f - pk[0][0] - pk[1][0]);
        pk[2][1] = clamp01(1.0f - pk[0][1] - pk[1][1]);
        pk[2][2] = clamp01(1.0f - pk[2][0] - pk[2][1]);
    }
}

}  // namespace cp3
