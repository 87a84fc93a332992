// The (cosZ, E) pairs of the mixed-precision propagation, shared by the single-grid and the
// batched kernels: propagate_rows_mixed<kRows> takes kRows cosine rows with the same number of
// crossed shells at one energy per lane.  Layer algebra and table layout: see
// propagate_pairs.hip.
//   L    : shells crossed (maxlayers of the rows)
//   gl[k]: layer lengths of row k in km, gl[k][0..L]
//   Q, M, R : the spectrum tables (float4 groups of v_1, v_2 / of the basis changes M_s /
//          double2 phase rates in half turns per km), row stride st
//   ie   : energy index (< ne_pad)
// Writes p[k][in][out] = P(in -> out) of row k.
#pragma once
#include "common.hpp"

namespace cp3 {

template <typename T> struct Cx { T re, im; };

// Forms of the running product (propagate_rows_mixed): the eigenbasis form with the basis
// changes read from the table (and the vacuum form for paths crossing no shell), or formed from
// the table's eigenvectors.
constexpr int kFormEigenTable = 1;
constexpr int kFormEigenLocal = 2;

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

// Cody-Waite style reduction of a phase of t half turns (phi = t pi) in FP64: q = rint(t),
// r = t - q in [-1/2, 1/2] is exact, so the phase reaches FP32 only as the remainder
// z = r pi in [-pi/2, pi/2]; there sin z and cos z - 1 are polynomials evaluated in FP32 by
// Horner's rule (sin z with its leading term z added last, exactly), and e^{i q pi} = (-1)^q is
// returned as a sign.  The coefficients are least-squares fits to sin and cos on [-pi/2, pi/2]
// with Lawson's iterative reweighting towards the minimax error: |sin z - s0| <= 4.6e-9 and
// |cos z - 1 - c0m1| <= 2.4e-10 before FP32 rounding (Taylor polynomials of the same degrees
// would leave 3.6e-6).  Folding pi into the coefficients (polynomials in r) would save the FP64
// multiply but puts pi's float rounding into the leading term of sin: |e^{i phi}|^2 - 1 then
// averages -3e-8 instead of 4e-9, a bias that grows linearly over the ~200 layer products of a
// core-crossing path on a 100-shell Earth.
__device__ inline __attribute__((always_inline)) void reduced_phase(double t, float& s0, float& c0m1, float& sign) {
#pragma clang fp contract(off)
    const double q = rint(t);
    const float z = float((t - q) * 3.141592653589793);
    const float z2 = z * z;
    float ps = __builtin_fmaf(z2, 2.600053676e-6f, -1.980661473e-4f);
    ps = __builtin_fmaf(z2, ps, 8.333017118e-3f);
    ps = __builtin_fmaf(z2, ps, -1.666665673e-1f);
    s0 = __builtin_fmaf(z * z2, ps, z);
    float pc = __builtin_fmaf(z2, -2.607709746e-7f, 2.476188638e-5f);
    pc = __builtin_fmaf(z2, pc, -1.388840377e-3f);
    pc = __builtin_fmaf(z2, pc, 4.166664183e-2f);
    pc = __builtin_fmaf(z2, pc, -0.5f);
    c0m1 = z2 * pc;
    // the parity of q is the last mantissa bit of q + 1.5 * 2^52: for |q| < 2^51 the sum lies in
    // [2^52, 2^53), where the last bit is worth 1, and the added even integer keeps the parity
    // (2^52 alone would not do for q < 0: the sum then falls below 2^52, where the last bit is
    // worth 1/2).  No 32-bit integer conversion, so phases up to 2^51 half turns are handled.
    const unsigned lo = unsigned(__builtin_bit_cast(unsigned long long, q + 6755399441055744.0));
    sign = __builtin_bit_cast(float, (lo << 31) | 0x3f800000u);
}

// e^{i phi} - 1 for a phase given in half turns, as (cos phi - 1, sin phi): with the sign,
// (c0m1, s0) for even q and (-c0m1 - 2, -s0) for odd q, each with one rounding.
__device__ inline void phase_factor_m1(double t, float& cm1, float& sn) {
    // every fused multiply-add is written out and nothing else may be fused, so that each
    // inlined copy (one row, two rows, batched) rounds identically
#pragma clang fp contract(off)
    float s0, c0m1, sign;
    reduced_phase(t, s0, c0m1, sign);
    cm1 = __builtin_fmaf(sign, c0m1, sign - 1.0f);
    sn = sign * s0;
}

// e^{i phi} for a phase in half turns: cos phi = sign (1 + (cos - 1)) and sin phi = sign sin,
// each with one rounding.
__device__ inline void phase_factor(double t, float& c, float& sn) {
#pragma clang fp contract(off)
    float s0, c0m1, sign;
    reduced_phase(t, s0, c0m1, sign);
    c = __builtin_fmaf(sign, c0m1, sign);
    sn = sign * s0;
}

// One shell's table entries at this lane's energy: the real eigenvectors v_1, v_2 and the two
// phase rates.  They depend on (shell, energy) only, so rows that cross the same shells share
// them.
struct LayerTable {
    float v[2][3];
    double2 r;
};

__device__ inline __attribute__((always_inline)) void load_table(const float4* __restrict__ Q,
                                                                const double2* __restrict__ R, std::size_t st,
                                                                unsigned off, int i, LayerTable& t) {
    // wave-uniform row addresses plus one 32-bit lane offset (both element types are 16 bytes),
    // so that the loads take the scalar-base form without 64-bit vector address arithmetic
    const char* q0 = reinterpret_cast<const char*>(Q + std::size_t(i) * kMixedQGroups * st);
    const float4 a = *reinterpret_cast<const float4*>(q0 + off);
    const float4 b = *reinterpret_cast<const float4*>(q0 + st * sizeof(float4) + off);
    t.v[0][0] = a.x; t.v[0][1] = a.y; t.v[0][2] = a.z;
    t.v[1][0] = a.w; t.v[1][1] = b.x; t.v[1][2] = b.y;
    t.r = *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(R + std::size_t(i) * st) + off);
}

// w_k = e^{i dphi_k} - 1 of a layer of length dist km
struct LayerPhases {
    float re[2], im[2];
};

__device__ inline __attribute__((always_inline)) LayerPhases layer_phases(const LayerTable& t, double dist) {
    LayerPhases w;
    phase_factor_m1(t.r.x * dist, w.re[0], w.im[0]);
    phase_factor_m1(t.r.y * dist, w.re[1], w.im[1]);
    return w;
}

// X <- S X for the layer amplitude S = I + w_1 v_1 v_1^T + w_2 v_2 v_2^T and a 3 x kCols
// complex X, as two rank-1 updates X += v_k (w_k (v_k^T X)) (applied one after the other:
// v_2^T v_1 = 0, so the second sees the same v_2^T X).  16 operations per column and k.
template <int kCols>
__device__ inline __attribute__((always_inline)) void apply_layer(const LayerTable& t, const LayerPhases& w,
                                                                  Cx<float> X[3][kCols]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const float* v = t.v[k];
#pragma unroll
        for (int j = 0; j < kCols; j++) {
            float cr = v[0] * X[0][j].re, ci = v[0] * X[0][j].im;
            cr = __builtin_fmaf(v[1], X[1][j].re, cr);
            ci = __builtin_fmaf(v[1], X[1][j].im, ci);
            cr = __builtin_fmaf(v[2], X[2][j].re, cr);
            ci = __builtin_fmaf(v[2], X[2][j].im, ci);
            float dr = w.re[k] * cr, di = w.re[k] * ci;
            dr = __builtin_fmaf(-w.im[k], ci, dr);
            di = __builtin_fmaf(w.im[k], cr, di);
#pragma unroll
            for (int a = 0; a < 3; a++) {
                X[a][j].re = __builtin_fmaf(v[a], dr, X[a][j].re);
                X[a][j].im = __builtin_fmaf(v[a], di, X[a][j].im);
            }
        }
    }
}

// One shell's basis change M_s = O_s^T O_{s-1} (rows) and phase rates at this lane's energy.
struct ShellBasis {
    float m[3][3];
    double2 r;
};

__device__ inline __attribute__((always_inline)) void load_basis(const float4* __restrict__ M,
                                                                const double2* __restrict__ R, std::size_t st,
                                                                unsigned off, int i, ShellBasis& b) {
    const char* m0 = reinterpret_cast<const char*>(M + std::size_t(i) * kMixedMGroups * st);
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const float4 a = *reinterpret_cast<const float4*>(m0 + r * st * sizeof(float4) + off);
        b.m[r][0] = a.x; b.m[r][1] = a.y; b.m[r][2] = a.z;
    }
    b.r = *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(R + std::size_t(i) * st) + off);
}

// The basis change formed from the stored eigenvectors instead (batched kernel, whose tables hold
// no basis changes): O_s = [v_1 x v_2, v_1, v_2] in float, M_s = O_s^T O_prev with O_prev the
// previous shell's eigenbasis (o[c] = column c), then O_prev <- O_s.  For s = 1, O_0 = I and
// M_1 = O_1^T.
__device__ inline __attribute__((always_inline)) void local_basis(const LayerTable& t, bool first, float oprev[3][3],
                                                                 ShellBasis& b) {
#pragma clang fp contract(off)
    const float* v1 = t.v[0];
    const float* v2 = t.v[1];
    float o[3][3];
    o[0][0] = __builtin_fmaf(v1[1], v2[2], -(v1[2] * v2[1]));
    o[0][1] = __builtin_fmaf(v1[2], v2[0], -(v1[0] * v2[2]));
    o[0][2] = __builtin_fmaf(v1[0], v2[1], -(v1[1] * v2[0]));
#pragma unroll
    for (int c = 0; c < 3; c++) {
        o[1][c] = v1[c];
        o[2][c] = v2[c];
    }
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            if (first) {
                b.m[r][c] = o[r][c];
            } else {
                float acc = o[r][0] * oprev[c][0];
                acc = __builtin_fmaf(o[r][1], oprev[c][1], acc);
                b.m[r][c] = __builtin_fmaf(o[r][2], oprev[c][2], acc);
            }
        }
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) oprev[r][c] = o[r][c];
    b.r = t.r;
}

// d_k = e^{i dphi_k} (k = 1, 2) of a layer of length dist km
struct LayerDiag {
    float c[2], s[2];
};

__device__ inline __attribute__((always_inline)) LayerDiag layer_diag(double2 r, double dist) {
    LayerDiag d;
    phase_factor(r.x * dist, d.c[0], d.s[0]);
    phase_factor(r.y * dist, d.c[1], d.s[1]);
    return d;
}

// The same, also adding both phases (half turns) to tsum: det D = e^{i pi (t_1 + t_2)}.
__device__ inline __attribute__((always_inline)) LayerDiag layer_diag(double2 r, double dist, double& tsum) {
    const double t1 = r.x * dist, t2 = r.y * dist;
    tsum += t1;
    tsum += t2;
    LayerDiag d;
    phase_factor(t1, d.c[0], d.s[0]);
    phase_factor(t2, d.c[1], d.s[1]);
    return d;
}

// Z <- D M Z for a real 3x3 M and D = diag(1, d_1, d_2): 3 x kCols complex Z.
template <int kCols>
__device__ inline __attribute__((always_inline)) void basis_step(const float m[3][3], const LayerDiag* d,
                                                                 Cx<float> Z[3][kCols]) {
#pragma clang fp contract(off)
    Cx<float> N[3][kCols];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int j = 0; j < kCols; j++) {
            float re = m[a][0] * Z[0][j].re, im = m[a][0] * Z[0][j].im;
            re = __builtin_fmaf(m[a][1], Z[1][j].re, re);
            im = __builtin_fmaf(m[a][1], Z[1][j].im, im);
            re = __builtin_fmaf(m[a][2], Z[2][j].re, re);
            im = __builtin_fmaf(m[a][2], Z[2][j].im, im);
            N[a][j] = Cx<float>{re, im};
        }
#pragma unroll
    for (int j = 0; j < kCols; j++) Z[0][j] = N[0][j];
#pragma unroll
    for (int a = 1; a < 3; a++)
#pragma unroll
        for (int j = 0; j < kCols; j++) {
            if (d) {
                const float c = d->c[a - 1], sn = d->s[a - 1];
                float re = c * N[a][j].re, im = c * N[a][j].im;
                re = __builtin_fmaf(-sn, N[a][j].im, re);
                im = __builtin_fmaf(sn, N[a][j].re, im);
                Z[a][j] = Cx<float>{re, im};
            } else {
                Z[a][j] = N[a][j];
            }
        }
}

// The unitary 3x3 Z from its first two columns and its determinant e^{i pi tsum}: columns 0 and
// 1 are normalised (their norms drift from 1 by the float rounding of the running product, the
// dominant error on long paths), and column 2 = det(Z) conj(c_0 x c_1), the cofactor identity of
// a unitary matrix (adj Z = det(Z) Z^dagger).
__device__ inline __attribute__((always_inline)) void complete_columns(const Cx<float> Z2[3][2], double tsum,
                                                                       Cx<float> Z[3][3]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int j = 0; j < 2; j++) {
        float n2 = Z2[0][j].re * Z2[0][j].re;
        n2 = __builtin_fmaf(Z2[0][j].im, Z2[0][j].im, n2);
#pragma unroll
        for (int a = 1; a < 3; a++) {
            n2 = __builtin_fmaf(Z2[a][j].re, Z2[a][j].re, n2);
            n2 = __builtin_fmaf(Z2[a][j].im, Z2[a][j].im, n2);
        }
        const float inv = __builtin_amdgcn_rsqf(n2);
#pragma unroll
        for (int a = 0; a < 3; a++) Z[a][j] = Cx<float>{Z2[a][j].re * inv, Z2[a][j].im * inv};
    }
    float dc, ds;
    phase_factor(tsum, dc, ds);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int b = (a + 1) % 3, c = (a + 2) % 3;
        Cx<float> x = cmul(Z[b][0], Z[c][1]);
        cmac(x, Cx<float>{-Z[c][0].re, -Z[c][0].im}, Z[b][1]);
        // det * conj(x)
        float re = dc * x.re, im = ds * x.re;
        re = __builtin_fmaf(ds, x.im, re);
        im = __builtin_fmaf(-dc, x.im, im);
        Z[a][2] = Cx<float>{re, im};
    }
}

// Atmosphere: G = S_0 W^dagger[:, (e, mu)] = W^dagger[:, (e, mu)] + sum_k v_k w_k z_k^T,
// z_k = W[(e, mu), :] v_k = (v_k0, W_mu1 v_k1 + W_mu2 v_k2) conjugated; t: slot 0's table entries.
__device__ inline __attribute__((always_inline)) void atmosphere_columns(const LayerTable& t, double dist, float4 wmu,
                                                                         Cx<float> g[3][2]) {
#pragma clang fp contract(off)
    const LayerPhases w = layer_phases(t, dist);
    g[0][0] = Cx<float>{1.0f, 0.0f}; g[1][0] = Cx<float>{0.0f, 0.0f}; g[2][0] = Cx<float>{0.0f, 0.0f};
    g[0][1] = Cx<float>{0.0f, 0.0f}; g[1][1] = Cx<float>{wmu.x, -wmu.y}; g[2][1] = Cx<float>{wmu.z, -wmu.w};
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
        const float* v = t.v[kk];
        const float zr = __builtin_fmaf(wmu.z, v[2], wmu.x * v[1]);
        const float zi = -__builtin_fmaf(wmu.w, v[2], wmu.y * v[1]);
        const float er = w.re[kk] * v[0], ei = w.im[kk] * v[0];
        float mr = w.re[kk] * zr, mi = w.re[kk] * zi;
        mr = __builtin_fmaf(-w.im[kk], zi, mr);
        mi = __builtin_fmaf(w.im[kk], zr, mi);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            g[a][0].re = __builtin_fmaf(v[a], er, g[a][0].re);
            g[a][0].im = __builtin_fmaf(v[a], ei, g[a][0].im);
            g[a][1].re = __builtin_fmaf(v[a], mr, g[a][1].re);
            g[a][1].im = __builtin_fmaf(v[a], mi, g[a][1].im);
        }
    }
}

// p[in][out] = |T[out][in]|^2 from the 2 x 2 block T (out, in in (e, mu)); the third row and
// column from unitarity
__device__ inline __attribute__((always_inline)) void block_probabilities(const Cx<float> T[2][2], float (&pk)[3][3]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int in = 0; in < 2; in++)
#pragma unroll
        for (int o = 0; o < 2; o++)
            pk[in][o] = __builtin_fmaf(T[o][in].re, T[o][in].re, T[o][in].im * T[o][in].im);
    pk[0][2] = clamp01(1.0f - pk[0][0] - pk[0][1]);
    pk[1][2] = clamp01(1.0f - pk[1][0] - pk[1][1]);
    pk[2][0] = clamp01(1.0f - pk[0][0] - pk[1][0]);
    pk[2][1] = clamp01(1.0f - pk[0][1] - pk[1][1]);
    pk[2][2] = clamp01(1.0f - pk[2][0] - pk[2][1]);
}

// The path's layers (physics.hpp:645-677): layer 0 (atmosphere), shells 1..L-1 on the way in,
// the innermost shell L, shells L-1..1 on the way out; total T = T_1..T_{L-1} T_L T_{L-1}..T_1 T_0.
// In the real basis every layer amplitude is W S_i W^dagger with S_i = O_i D_i O_i^T complex
// symmetric (O_i the real eigenbasis, D_i = diag(1, e^{i dphi_1}, e^{i dphi_2}) with the global
// phase dropped: it cancels in |T|^2), so with Y = S_{L-1} ... S_1 (the way in) the way out
// S_1 ... S_{L-1} is Y^T, and
//   T = W Y^T S_L Y G,   G = S_0 W^dagger[:, (e, mu)].
// Only the flavour rows and columns e, mu of T are needed (unitarity gives the rest).  For L >= 2
// the running product is carried in the shells' eigenbases (kFormEigenTable, single-grid
// kernels; kFormEigenLocal, batched kernel):
//   Y = O_{L-1} Z with Z = D_{L-1} M_{L-1} ... D_1 M_1, M_s = O_s^T O_{s-1} (O_0 = I), read from
//   the table's basis changes or formed from its eigenvectors (local_basis).  Z is unitary with
//   det Z = prod_s det D_s (det M_s = 1), so only its first two columns are carried: one shell
//   costs a real 3x3 times complex 3x2 product, two complex row scalings and two FP64 phase sums;
//   after Z' = M_L Z the columns are normalised and the third is det(Z') conj(c_0 x c_1).  With
//   Z' = O_L^T Y,
//     T[x][y] = sum_c (Z' a_x)_c D_L,c (Z' g_y)_c,
//   a_e = e_0, a_mu = row mu of W (as a column), g_y = G[:, y].
// L <= 1: T = W S_1 G (S_1 as two rank-1 updates of G) or W G; in the single-grid kernel a path
// crossing no shell takes the vacuum form T = U D_0 U^dagger with host constants U_xk conj(U_yk).
//
// kRows cosine rows with the same L are propagated together at this lane's energy: they cross
// the same shells, so each shell's table entries are loaded once for all of them (only the
// layer lengths, hence the phases, differ), which halves the table reads per pair for kRows = 2.
// wmu = row mu of W (OscParams::Wmu); row e of W is e_0^T.  L is wave-uniform, so every branch
// here is a scalar one.
template <int kRows, int kForm>
__device__ inline void propagate_rows_mixed(int L, const double* const* gl, const float4* __restrict__ Q,
                                            const float4* __restrict__ M, const double2* __restrict__ R,
                                            std::size_t st, int ie, float4 wmu, const VacuumBlock* vac,
                                            float (*p)[3][3]) {
#pragma clang fp contract(off)
    static_assert(sizeof(float4) == 16 && sizeof(double2) == 16, "table entries are 16 bytes");
    const unsigned off = unsigned(ie) * 16u;
    const Cx<float> w1{wmu.x, wmu.y}, w2{wmu.z, wmu.w};
    LayerTable t;
    Cx<float> G[kRows][3][2];   // columns e, mu of S_0 W^dagger
    if (kForm == kFormEigenTable && L == 0) {
        // one vacuum layer: T[x][y] = delta_xy + sum_{k=1,2} w_k U_xk conj(U_yk), w_k = e^{i dphi_k} - 1
        // (W S_0 W^dagger = U D U^dagger in vacuum); only the phase rates are read from the table
        const double2 r = *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(R) + off);
#pragma unroll
        for (int k = 0; k < kRows; k++) {
            LayerPhases w;
            phase_factor_m1(r.x * gl[k][0], w.re[0], w.im[0]);
            phase_factor_m1(r.y * gl[k][0], w.re[1], w.im[1]);
            Cx<float> T[2][2];
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++) {
                    const float* c1 = vac->c + (x * 2 + y) * 2;
                    const float* c2 = c1 + 8;
                    float re = x == y ? 1.0f : 0.0f, im = 0.0f;
                    re = __builtin_fmaf(w.re[0], c1[0], re);
                    re = __builtin_fmaf(-w.im[0], c1[1], re);
                    re = __builtin_fmaf(w.re[1], c2[0], re);
                    re = __builtin_fmaf(-w.im[1], c2[1], re);
                    im = __builtin_fmaf(w.re[0], c1[1], im);
                    im = __builtin_fmaf(w.im[0], c1[0], im);
                    im = __builtin_fmaf(w.re[1], c2[1], im);
                    im = __builtin_fmaf(w.im[1], c2[0], im);
                    T[x][y] = Cx<float>{re, im};
                }
            block_probabilities(T, p[k]);
        }
        return;
    }
    if (L < 2) {
        load_table(Q, R, st, off, 0, t);
#pragma unroll
        for (int k = 0; k < kRows; k++) atmosphere_columns(t, gl[k][0], wmu, G[k]);
        if (L == 1) {
            load_table(Q, R, st, off, 1, t);
#pragma unroll
            for (int k = 0; k < kRows; k++) apply_layer<2>(t, layer_phases(t, gl[k][1]), G[k]);
        }
#pragma unroll
        for (int k = 0; k < kRows; k++) {
            Cx<float> T[2][2];
#pragma unroll
            for (int b = 0; b < 2; b++) {
                T[0][b] = G[k][0][b];
                Cx<float> m = cmul(w1, G[k][1][b]);
                cmac(m, w2, G[k][2][b]);
                T[1][b] = m;
            }
            block_probabilities(T, p[k]);
        }
        return;
    }
    {
        ShellBasis b;
        Cx<float> Z2[kRows][3][2];   // columns 0, 1 of Z
        double tsum[kRows];          // det Z = e^{i pi tsum}
        float oprev[3][3];           // kFormEigenLocal: the previous shell's eigenbasis
        LayerTable tn;               // kFormEigenLocal: the next shell's entries, loaded one shell ahead
        auto basis = [&](int i) {
            if constexpr (kForm == kFormEigenTable) {
                load_basis(M, R, st, off, i, b);
            } else {
                if (i == 1) load_table(Q, R, st, off, 1, tn);
                t = tn;
                // in flight while this shell is computed (the batched kernel's table slices miss
                // L2 more often than the single grid's)
                if (i < L) load_table(Q, R, st, off, i + 1, tn);
                local_basis(t, i == 1, oprev, b);
            }
        };
        basis(1);
#pragma unroll
        for (int k = 0; k < kRows; k++) {           // Z = D_1 M_1
            tsum[k] = 0.0;
            const LayerDiag d = layer_diag(b.r, gl[k][1], tsum[k]);
#pragma unroll
            for (int j = 0; j < 2; j++) {
                Z2[k][0][j] = Cx<float>{b.m[0][j], 0.0f};
#pragma unroll
                for (int a = 1; a < 3; a++) Z2[k][a][j] = Cx<float>{d.c[a - 1] * b.m[a][j], d.s[a - 1] * b.m[a][j]};
            }
        }
        auto shell = [&](int i) {
            basis(i);
#pragma unroll
            for (int k = 0; k < kRows; k++) {
                const LayerDiag d = layer_diag(b.r, gl[k][i], tsum[k]);
                basis_step<2>(b.m, &d, Z2[k]);
            }
        };
        // one row: unrolled twice for independent work across shells; two rows carry enough
        // independent work as they are
        if (kRows == 1) {
#pragma unroll 2
            for (int i = 2; i < L; i++) shell(i);
        } else {
#pragma unroll 1
            for (int i = 2; i < L; i++) shell(i);
        }
        basis(L);
        LayerDiag dL[kRows];
#pragma unroll
        for (int k = 0; k < kRows; k++) {
            basis_step<2>(b.m, nullptr, Z2[k]);     // Z' = M_L Z, columns 0, 1
            dL[k] = layer_diag(b.r, gl[k][L]);
        }
        load_table(Q, R, st, off, 0, t);
#pragma unroll
        for (int k = 0; k < kRows; k++) {
            Cx<float> Z[3][3];                      // Z'
            complete_columns(Z2[k], tsum[k], Z);
            atmosphere_columns(t, gl[k][0], wmu, G[k]);
            // h_x = D_L Z' a_x for x = e, mu, and u_y = Z' g_y
            Cx<float> h[2][3], u[3][2];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                h[0][c] = Z[c][0];
                h[1][c] = cmul(w1, Z[c][1]);
                cmac(h[1][c], w2, Z[c][2]);
#pragma unroll
                for (int y = 0; y < 2; y++) {
                    Cx<float> acc = cmul(Z[c][0], G[k][0][y]);
                    cmac(acc, Z[c][1], G[k][1][y]);
                    cmac(acc, Z[c][2], G[k][2][y]);
                    u[c][y] = acc;
                }
            }
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int c = 1; c < 3; c++) h[x][c] = cmul(Cx<float>{dL[k].c[c - 1], dL[k].s[c - 1]}, h[x][c]);
            Cx<float> T[2][2];
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++) {
                    Cx<float> acc = cmul(h[x][0], u[0][y]);
                    cmac(acc, h[x][1], u[1][y]);
                    cmac(acc, h[x][2], u[2][y]);
                    T[x][y] = acc;
                }
            block_probabilities(T, p[k]);
        }
    }
}

}  // namespace cp3
