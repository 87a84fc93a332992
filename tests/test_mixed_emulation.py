"""CPU emulation of the mixed-precision algebra of the GPU kernels (pair_mixed.hpp) against the
reference's golden vectors: pins the precision design itself, independently of a GPU.

What is emulated, stage by stage (numpy, vectorised over the energies of a cosine row):
  * the spectrum per (shell, energy) in FP64 -- eigenvalues and flavour-basis projectors
    Q_k = U P_k U^dagger of the layer Hamiltonian, here from numpy.linalg.eigh instead of the
    reference's cubic formula (equal to ~1e-16), Q_1 and Q_2 rounded to float32 as stored;
  * the phase differences dphi_k = -2.534 (lambda_k - lambda_0) L / E in quarter turns, reduced in
    FP64 (q = rint(t), r = t - q), the remainder z = r pi / 2 rounded to float32, sin z and
    cos z - 1 by the kernel's Taylor polynomials in float32, rotated by the quadrant q mod 4;
  * A = I + w_1 Q_1 + w_2 Q_2, the running products (columns 0-1 forward, rows 0-1 mantle) and
    S = mantle x forward in complex64; the upper-left 2x2 block of |S|^2 and the other five
    probabilities from unitarity, clamped to [0, 1].
The bar is the mixed kernels' 1e-5 (max |dP| over all nine channels).
"""
import numpy as np
import pytest

import oracle
from conftest import golden_names, load_golden
from cudaprob3_amd.propagator import Propagator

MIXED_TOL = 1e-5
KGF, LOE = 1.52588e-4, 2.534
F32 = np.float32


def spectrum(U18, dm9, E, rho_eff, antinu):
    """Eigenvalues (ascending) and flavour-basis projectors of the layer Hamiltonian at energies E
    (FP64): mass-basis H = diag(0, DM(1,0), DM(2,0)) + a w w^dagger, w = conj(U_e.), a = +-GF E rho."""
    U = (U18[0::2] + 1j * U18[1::2]).reshape(3, 3)
    a = (-1.0 if antinu else 1.0) * KGF * E * rho_eff
    w = np.conj(U[0, :])
    H = np.diag([0.0, dm9[3], dm9[6]]).astype(complex)[None] + a[:, None, None] * np.outer(w, np.conj(w))[None]
    lam, V = np.linalg.eigh(H)
    Q = np.einsum("ij,ejk,elk,ml->ekim", U, V, np.conj(V), np.conj(U))   # [e][k] U v_k v_k^dagger U^dagger
    return lam, Q


def phase_factor_m1(t):
    """e^{i phi} - 1 for phases t in quarter turns: FP64 reduction, float32 polynomials."""
    q = np.rint(t)
    z = ((t - q) * 1.5707963267948966).astype(F32)
    z2 = z * z
    ps = z2 * F32(2.75573192e-6) + F32(-1.98412698e-4)
    ps = z2 * ps + F32(8.33333333e-3)
    ps = z2 * ps + F32(-1.66666667e-1)
    s0 = (z * z2) * ps + z
    pc = z2 * F32(-2.75573192e-7) + F32(2.48015873e-5)
    pc = z2 * pc + F32(-1.38888889e-3)
    pc = z2 * pc + F32(4.16666667e-2)
    pc = z2 * pc + F32(-0.5)
    c0m1 = z2 * pc
    quad = np.asarray(q, np.int64) & 3
    one, two = F32(1), F32(2)
    cm1 = np.select([quad == 0, quad == 1, quad == 2], [c0m1, -s0 - one, -c0m1 - two], s0 - one)
    sn = np.select([quad == 0, quad == 1, quad == 2], [s0, c0m1 + one, -s0], -(c0m1 + one))
    return (cm1.astype(F32) + 1j * sn.astype(F32)).astype(np.complex64)


def emulate(g, phase=phase_factor_m1):
    p = Propagator(len(g["cosines"]), len(g["energies"]))
    p.setEnergyList(g["energies"])
    p.setCosineList(g["cosines"])
    p.setDensity(g["radii"], g["rhos"])
    p.setProductionHeight(float(g["h_km"]))
    p.setMNSMatrix(*g["params"][:4])
    p.setNeutrinoMasses(*g["params"][4:])
    antinu = int(g["antinu"])
    E = np.asarray(p.energyList, np.float64)
    n_e = len(E)
    # per shell slot: spectrum at every energy (slot 0 = the vacuum atmosphere)
    tables = []
    for slot in range(len(p.radii) + 1):
        rho = 0.0 if slot == 0 else float(p.rhos[slot - 1]) * 0.5
        lam, Q = spectrum(np.asarray(p.Mix_U), np.asarray(p.dm), E, np.full(n_e, rho), antinu)
        rate = [(-LOE * (lam[:, k] - lam[:, 0]) / E) / (np.pi / 2) for k in (1, 2)]   # quarter turns per km
        tables.append((rate, Q[:, 1].astype(np.complex64), Q[:, 2].astype(np.complex64)))
    out = np.zeros((len(p.cosineList), n_e, 9))
    eye = np.eye(3, dtype=np.complex64)[None]
    for ic, cz in enumerate(p.cosineList):
        L = int(p.maxlayers[ic])
        dist = oracle.layer_distances(p.radii, L, cz, float(p.ProductionHeightinCentimeter)) / 1e5
        F = M = None
        for i in range(L + 1):
            rate, Q1, Q2 = tables[i]
            w1 = phase(rate[0] * dist[i])
            w2 = phase(rate[1] * dist[i])
            A = (eye + w1[:, None, None] * Q1 + w2[:, None, None] * Q2).astype(np.complex64)
            if i == 0:
                F = A[:, :, :2]
            else:
                F = (A @ F).astype(np.complex64)
                if i < L:
                    M = A[:, :2, :] if i == 1 else (M @ A).astype(np.complex64)
        S = (M @ F).astype(np.complex64) if L >= 2 else F[:, :2, :]
        P2 = (np.abs(S) ** 2).astype(F32)                      # P2[e, out, in]
        P = np.zeros((n_e, 3, 3), F32)                          # P[e, in, out]
        P[:, :2, :2] = np.transpose(P2, (0, 2, 1))
        P[:, 0, 2] = np.clip(1 - P[:, 0, 0] - P[:, 0, 1], 0, 1)
        P[:, 1, 2] = np.clip(1 - P[:, 1, 0] - P[:, 1, 1], 0, 1)
        P[:, 2, 0] = np.clip(1 - P[:, 0, 0] - P[:, 1, 0], 0, 1)
        P[:, 2, 1] = np.clip(1 - P[:, 0, 1] - P[:, 1, 1], 0, 1)
        P[:, 2, 2] = np.clip(1 - P[:, 2, 0] - P[:, 2, 1], 0, 1)
        out[ic] = P.reshape(n_e, 9)
    return out


@pytest.mark.parametrize("name", golden_names())
def test_mixed_algebra_meets_tolerance(name):
    g = load_golden(name)
    P = emulate(g)
    ref = g["P"]
    err = np.abs(P - ref).max()
    print(f"{name}: emulated mixed algebra max|dP| {err:.2e}")
    assert err <= MIXED_TOL


def test_emulation_detects_a_wrong_quadrant():
    """The check is sharp: a quarter turn added to every phase difference (the quadrant rotation
    off by one) breaks the bar on a core-crossing fixture."""
    g = load_golden("prem59_nu")
    good = emulate(g)
    bad = emulate(g, phase=lambda t: phase_factor_m1(t + 1.0))
    assert np.abs(good - g["P"]).max() <= MIXED_TOL
    assert np.abs(bad - g["P"]).max() > 1e-2
