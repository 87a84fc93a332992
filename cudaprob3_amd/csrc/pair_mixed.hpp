// One (cosZ, E) pair of the mixed-precision propagation, shared by the single-grid and the
// batched kernels.  Layer algebra and table layout: see propagate_pairs.hip.
//   L   : shells crossed (maxlayers of the row)
//   gl  : layer lengths of the row in km, gl[0..L]
//   Q, R: the energy tile's spectrum table (float4 groups / double2 phase rates), row stride st
//   ie  : energy index (< ne_pad)
// Writes p[in][out] = P(in -> out).
#pragma once
#include "common.hpp"

namespace cp3 {

template <typename T> struct Cx { T re, im; };

__device__ inline float clamp01(float x) { return fminf(1.0f, fmaxf(0.0f, x)); }

__device__ inline void propagate_pair_mixed(int L, const double* __restrict__ gl, const float4* __restrict__ Q,
                                            const double2* __restrict__ R, std::size_t st, int ie,
                                            float p[3][3]) {
    Cx<float> F[3][2];   // columns 0-1 of T_i ... T_1 T_atm
    Cx<float> M[2][3];   // rows 0-1 of T_1 ... T_{i}
    for (int i = 0; i <= L; i++) {
        const double dist = gl[i];
        const float4* qp = Q + std::size_t(i) * 5 * st + ie;
        float q[20];
#pragma unroll
        for (int grp = 0; grp < 5; grp++) {
            const float4 v = qp[grp * st];
            q[4 * grp] = v.x; q[4 * grp + 1] = v.y; q[4 * grp + 2] = v.z; q[4 * grp + 3] = v.w;
        }
        const double2 r = R[std::size_t(i) * st + ie];
        float wr[2], wi[2];
        {
            double t1 = r.x * dist, t2 = r.y * dist;
            t1 -= rint(t1);
            t2 -= rint(t2);
            const float x1 = float(t1), x2 = float(t2);
            wr[0] = __builtin_amdgcn_cosf(x1) - 1.0f;
            wi[0] = __builtin_amdgcn_sinf(x1);
            wr[1] = __builtin_amdgcn_cosf(x2) - 1.0f;
            wi[1] = __builtin_amdgcn_sinf(x2);
        }
        // A = I + w_1 Q_1 + w_2 Q_2
        Cx<float> A[3][3];
#pragma unroll
        for (int d = 0; d < 3; d++)
            A[d][d] = Cx<float>{1.0f + wr[0] * q[d] + wr[1] * q[9 + d], wi[0] * q[d] + wi[1] * q[9 + d]};
        const int idx[3][2] = {{0, 1}, {0, 2}, {1, 2}};
#pragma unroll
        for (int w = 0; w < 3; w++) {
            float ure = 0.0f, uim = 0.0f, lre = 0.0f, lim = 0.0f;
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const float re = q[9 * k + 3 + 2 * w], im = q[9 * k + 4 + 2 * w];
                ure += wr[k] * re - wi[k] * im;
                uim += wr[k] * im + wi[k] * re;
                lre += wr[k] * re + wi[k] * im;
                lim += wi[k] * re - wr[k] * im;
            }
            A[idx[w][0]][idx[w][1]] = Cx<float>{ure, uim};
            A[idx[w][1]][idx[w][0]] = Cx<float>{lre, lim};
        }
        if (i == 0) {
#pragma unroll
            for (int x = 0; x < 3; x++)
#pragma unroll
                for (int y = 0; y < 2; y++) F[x][y] = A[x][y];
        } else {
            Cx<float> G[3][2];
#pragma unroll
            for (int x = 0; x < 3; x++)
#pragma unroll
                for (int y = 0; y < 2; y++) {
                    float re = 0.0f, im = 0.0f;
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        re += A[x][k].re * F[k][y].re - A[x][k].im * F[k][y].im;
                        im += A[x][k].re * F[k][y].im + A[x][k].im * F[k][y].re;
                    }
                    G[x][y] = Cx<float>{re, im};
                }
#pragma unroll
            for (int x = 0; x < 3; x++)
#pragma unroll
                for (int y = 0; y < 2; y++) F[x][y] = G[x][y];
            if (i < L) {
                if (i == 1) {
#pragma unroll
                    for (int x = 0; x < 2; x++)
#pragma unroll
                        for (int y = 0; y < 3; y++) M[x][y] = A[x][y];
                } else {
                    Cx<float> N[2][3];
#pragma unroll
                    for (int x = 0; x < 2; x++)
#pragma unroll
                        for (int y = 0; y < 3; y++) {
                            float re = 0.0f, im = 0.0f;
#pragma unroll
                            for (int k = 0; k < 3; k++) {
                                re += M[x][k].re * A[k][y].re - M[x][k].im * A[k][y].im;
                                im += M[x][k].re * A[k][y].im + M[x][k].im * A[k][y].re;
                            }
                            N[x][y] = Cx<float>{re, im};
                        }
#pragma unroll
                    for (int x = 0; x < 2; x++)
#pragma unroll
                        for (int y = 0; y < 3; y++) M[x][y] = N[x][y];
                }
            }
        }
    }
    // upper-left 2x2 block S[out][in] of the total amplitude
    Cx<float> S[2][2];
    if (L >= 2) {
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 2; y++) {
                float re = 0.0f, im = 0.0f;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    re += M[x][k].re * F[k][y].re - M[x][k].im * F[k][y].im;
                    im += M[x][k].re * F[k][y].im + M[x][k].im * F[k][y].re;
                }
                S[x][y] = Cx<float>{re, im};
            }
    } else {
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 2; y++) S[x][y] = F[x][y];
    }
    // p[in][out] = |S[out][in]|^2; the third row and column from unitarity
#pragma unroll
    for (int in = 0; in < 2; in++)
#pragma unroll
        for (int o = 0; o < 2; o++) p[in][o] = S[o][in].re * S[o][in].re + S[o][in].im * S[o][in].im;
    p[0][2] = clamp01(1.0f - p[0][0] - p[0][1]);
    p[1][2] = clamp01(1.0f - p[1][0] - p[1][1]);
    p[2][0] = clamp01(1.0f - p[0][0] - p[1][0]);
    p[2][1] = clamp01(1.0f - p[0][1] - p[1][1]);
    p[2][2] = clamp01(1.0f - p[2][0] - p[2][1]);
}

}  // namespace cp3
