"""CPU emulation of the mixed-precision algebra of the GPU kernels (pair_mixed.hpp) against the
reference's golden vectors: pins the precision design itself, independently of a GPU.  Two forms
of the running product are emulated: rank-1 updates (emulate; the kernels' form for rows crossing
at most one shell, and the earlier form of every row) and the eigenbasis form of rows crossing two
or more shells, with the table's basis changes (single-grid kernels) or with basis changes formed
from the float eigenvectors (batched kernel) (emulate_eigenbasis).

What is emulated, stage by stage (numpy, vectorised over the energies of a cosine row):
  * the real basis (host_util.hpp real_basis): the unitary W = diag(1, W') that makes the vacuum
    flavour Hamiltonian U diag(0, DM(1,0), DM(2,0)) U^dagger real symmetric (Hr) while fixing
    the electron flavour, built from U as the library builds it;
  * the spectrum per (shell, energy) in FP64 -- eigenvalues and real unit eigenvectors v_k of
    Hr + a e_0 e_0^T, here from numpy.linalg.eigh instead of the reference's cubic formula and
    eq. 11 projectors (equal to ~1e-16), v_1 and v_2 rounded to float32 as stored;
  * the phase differences dphi_k = -2.534 (lambda_k - lambda_0) L / E in half turns, reduced in
    FP64 (q = rint(t), r = t - q), the remainder z = r pi rounded to float32, sin z and
    cos z - 1 by the kernel's fitted polynomials in float32, sign-flipped for odd q;
  * the running product Y = S_{L-1} ... S_1 of the way in by rank-1 updates
    Y += v_k (w_k v_k^T Y), the way out as Y^T, the atmosphere as S_0 W^dagger[:, e mu] and the
    2x2 block (W Y^T)[e mu, :] S_L Y S_0 W^dagger[:, e mu] in complex64; the upper-left 2x2 block
    of |T|^2 and the other five probabilities from unitarity, clamped to [0, 1].
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


def real_basis(U, dm9):
    """(Hr, W): W = diag(1, W') unitary with W^dagger M W real symmetric for the vacuum part
    M = U diag(0, DM(1,0), DM(2,0)) U^dagger, W' = [a, b]: a along (M_10, M_20), b orthogonal to
    it with the phase that makes (W^dagger M W)_12 real."""
    M = sum(dm9[3 * n] * np.outer(U[:, n], np.conj(U[:, n])) for n in (1, 2))
    x = M[1:, 0]
    nx = np.linalg.norm(x)
    a = x / nx if nx > 0 else np.array([1.0, 0.0], complex)
    b = np.array([-np.conj(a[1]), np.conj(a[0])])
    m = np.conj(a) @ M[1:, 1:] @ b
    if abs(m) > 0:
        b = b * np.conj(m) / abs(m)
    W = np.eye(3, dtype=complex)
    W[1:, 1] = a
    W[1:, 2] = b
    H = np.conj(W.T) @ M @ W
    return H, W


def spectrum(Hr, E, rho_eff, antinu):
    """Eigenvalues (ascending) and real eigenvectors [e][component][k] of Hr + a e_0 e_0^T at
    energies E (FP64), a = +-GF E rho."""
    a = (-1.0 if antinu else 1.0) * KGF * E * rho_eff
    H = np.repeat(Hr.real[None], len(E), axis=0)
    H[:, 0, 0] += a
    return np.linalg.eigh(H)


SIN_C = [F32(c) for c in (-1.666665673e-1, 8.333017118e-3, -1.980661473e-4, 2.600053676e-6)]
COS_C = [F32(c) for c in (-0.5, 4.166664183e-2, -1.388840377e-3, 2.476188638e-5, -2.607709746e-7)]


def sign_bit(q):
    """The kernels' parity of the integer-valued FP64 q: the last mantissa bit of q + 1.5 * 2^52."""
    return (np.asarray(q, np.float64) + 6755399441055744.0).view(np.uint64) & 1


def test_sign_bit_is_the_parity():
    """The parity trick of phase_factor_m1 (pair_mixed.hpp) over negative and positive q up to
    2^51, against integer parity; the 2^52 variant fails for negative odd q."""
    rng = np.random.default_rng(3)
    q = np.concatenate([np.arange(-5000, 5000), rng.integers(-2 ** 51 + 1, 2 ** 51, 200_000),
                        [2 ** 51 - 1, -(2 ** 51) + 1, 2 ** 31, -(2 ** 31) - 1]]).astype(np.float64)
    assert np.array_equal(sign_bit(q), q.astype(np.int64) & 1)
    wrong = (q + 4503599627370496.0).view(np.uint64) & 1
    assert not np.array_equal(wrong, q.astype(np.int64) & 1)


def reduced_phase(t):
    """(sin z, cos z - 1, (-1)^q) of phases t in half turns as the kernels form them: FP64
    reduction q = rint(t), z = (t - q) pi rounded to float32, float32 polynomials in z."""
    q = np.rint(t)
    z = ((t - q) * np.pi).astype(F32)
    z2 = z * z
    ps = SIN_C[-1]
    for c in SIN_C[-2::-1]:
        ps = z2 * ps + c
    s0 = (z * z2) * ps + z
    pc = COS_C[-1]
    for c in COS_C[-2::-1]:
        pc = z2 * pc + c
    c0m1 = z2 * pc
    sign = np.where(sign_bit(q), F32(-1), F32(1)).astype(F32)
    return s0, c0m1, sign


def phase_factor_m1(t):
    """e^{i phi} - 1 for phases t in half turns: FP64 reduction, float32 polynomials, sign (-1)^q."""
    s0, c0m1, sign = reduced_phase(t)
    cm1 = sign * c0m1 + (sign - F32(1))
    sn = sign * s0
    return (cm1.astype(F32) + 1j * sn.astype(F32)).astype(np.complex64)


def test_phase_factor_against_the_exponential():
    """The kernel's phase polynomials against e^{i pi t} - 1 (numpy, FP64) for phases up to 10^4
    half turns and small ones: absolute error at the FP32 rounding level (2.5e-7)."""
    rng = np.random.default_rng(7)
    t = np.concatenate([rng.uniform(-1e4, 1e4, 400_000), rng.uniform(-2, 2, 400_000), [0.0, 0.5, -0.5, 1.0]])
    err = np.abs(phase_factor_m1(t).astype(np.complex128) - (np.exp(1j * np.pi * t) - 1))
    assert err.max() < 2.5e-7


def apply_layer(X, v, w):
    """X <- S X for S = I + sum_k w_k v_k v_k^T: X [e][3][cols] complex64, v [k][e][3] float32."""
    for k in range(2):
        c = np.einsum("ea,eaj->ej", v[k], X).astype(np.complex64)
        d = (w[k][:, None] * c).astype(np.complex64)
        X = (X + v[k][:, :, None] * d[:, None, :]).astype(np.complex64)
    return X


def emulate(g, phase=phase_factor_m1, way_out=lambda Y: np.transpose(Y, (0, 2, 1))):
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
    U18 = np.asarray(p.Mix_U)
    Hr, W = real_basis((U18[0::2] + 1j * U18[1::2]).reshape(3, 3), np.asarray(p.dm))
    Wc = np.conj(W[:2, :].T).astype(np.complex64)           # W^dagger[:, (e, mu)]
    Wr = W[:2, :].astype(np.complex64)                      # W[(e, mu), :]
    # per shell slot: phase rates and eigenvectors at every energy (slot 0 = the vacuum atmosphere)
    tables = []
    for slot in range(len(p.radii) + 1):
        rho = 0.0 if slot == 0 else float(p.rhos[slot - 1]) * 0.5
        lam, V = spectrum(Hr, E, np.full(n_e, rho), antinu)
        rate = [(-LOE * (lam[:, k] - lam[:, 0]) / E) / np.pi for k in (1, 2)]   # half turns per km
        tables.append((rate, [V[:, :, k].astype(F32) for k in (1, 2)]))
    out = np.zeros((len(p.cosineList), n_e, 9))
    for ic, cz in enumerate(p.cosineList):
        L = int(p.maxlayers[ic])
        dist = oracle.layer_distances(p.radii, L, cz, float(p.ProductionHeightinCentimeter)) / 1e5

        def layer(i):
            rate, v = tables[i]
            return v, [phase(rate[0] * dist[i]), phase(rate[1] * dist[i])]

        if L >= 2:
            v, w = layer(1)
            Y = apply_layer(np.repeat(np.eye(3, dtype=np.complex64)[None], n_e, axis=0), v, w)
            for i in range(2, L):
                Y = apply_layer(Y, *layer(i))
        G = apply_layer(np.repeat(Wc[None], n_e, axis=0), *layer(0))
        if L >= 2:
            G = (Y @ G).astype(np.complex64)
        if L >= 1:
            G = apply_layer(G, *layer(L))
        Hrows = (Wr[None] @ way_out(Y)).astype(np.complex64) if L >= 2 else Wr[None]
        T = (Hrows @ G).astype(np.complex64)                  # T[e, out, in]
        P2 = (np.abs(T) ** 2).astype(F32)
        P = np.zeros((n_e, 3, 3), F32)                          # P[e, in, out]
        P[:, :2, :2] = np.transpose(P2, (0, 2, 1))
        P[:, 0, 2] = np.clip(1 - P[:, 0, 0] - P[:, 0, 1], 0, 1)
        P[:, 1, 2] = np.clip(1 - P[:, 1, 0] - P[:, 1, 1], 0, 1)
        P[:, 2, 0] = np.clip(1 - P[:, 0, 0] - P[:, 1, 0], 0, 1)
        P[:, 2, 1] = np.clip(1 - P[:, 0, 1] - P[:, 1, 1], 0, 1)
        P[:, 2, 2] = np.clip(1 - P[:, 2, 0] - P[:, 2, 1], 0, 1)
        out[ic] = P.reshape(n_e, 9)
    return out


def test_real_basis_is_real_and_fixes_the_electron_row():
    """W is unitary, keeps e_0, and W^dagger M W is real for a CP-violating U; the same holds for
    a U with arbitrary row and column phases (no phase convention assumed)."""
    g = load_golden("prem12_io_dcp")
    p = Propagator(1, 1)
    p.setMNSMatrix(*g["params"][:4])
    p.setNeutrinoMasses(*g["params"][4:])
    U18 = np.asarray(p.Mix_U)
    U = (U18[0::2] + 1j * U18[1::2]).reshape(3, 3)
    rng = np.random.default_rng(1)
    for Ux in (U, np.diag(np.exp(1j * rng.uniform(0, 6, 3))) @ U @ np.diag(np.exp(1j * rng.uniform(0, 6, 3)))):
        Hr, W = real_basis(Ux, np.asarray(p.dm))
        assert np.abs(np.conj(W.T) @ W - np.eye(3)).max() < 1e-15
        assert np.abs(W[0] - [1, 0, 0]).max() == 0 and np.abs(W[:, 0] - [1, 0, 0]).max() == 0
        M = sum(np.asarray(p.dm)[3 * n] * np.outer(Ux[:, n], np.conj(Ux[:, n])) for n in (1, 2))
        scale = np.abs(M).max()
        assert np.abs(Hr.imag).max() < 1e-14 * scale
        assert np.abs(W @ Hr @ np.conj(W.T) - M).max() < 1e-14 * scale


@pytest.mark.parametrize("name", golden_names())
def test_mixed_algebra_meets_tolerance(name):
    g = load_golden(name)
    P = emulate(g)
    ref = g["P"]
    err = np.abs(P - ref).max()
    print(f"{name}: emulated mixed algebra max|dP| {err:.2e}")
    assert err <= MIXED_TOL


def test_emulation_detects_a_wrong_sign():
    """The check is sharp: half a turn added to every phase difference (the sign (-1)^q taken
    from the wrong parity) breaks the bar on a core-crossing fixture."""
    g = load_golden("prem59_nu")
    good = emulate(g)
    bad = emulate(g, phase=lambda t: phase_factor_m1(t + 1.0))
    assert np.abs(good - g["P"]).max() <= MIXED_TOL
    assert np.abs(bad - g["P"]).max() > 1e-2


def test_emulation_detects_a_wrong_way_out():
    """The way out is the transpose of the way in (S_1 ... S_{L-1}); taking the product of the way
    in itself (S_{L-1} ... S_1, the shells in the wrong order) breaks the bar."""
    g = load_golden("prem59_nu")
    bad = emulate(g, way_out=lambda Y: Y)
    assert np.abs(bad - g["P"]).max() > 1e-2


def phase_factor(t):
    """e^{i phi} for phases t in half turns, as the eigenbasis path forms it: the same FP64
    reduction and float32 polynomials, cos = sign (1 + (cos - 1)) and sin = sign sin."""
    s0, c0m1, sign = reduced_phase(t)
    return ((sign * c0m1 + sign).astype(F32) + 1j * (sign * s0).astype(F32)).astype(np.complex64)


def eigenbasis_tables(Hr, E, rhos, antinu):
    """Per shell slot: phase rates, v_1, v_2 (float32) and the basis change M_s = O_s^T O_{s-1}
    (float32 from FP64), O_s = [v_1 x v_2, v_1, v_2] (det +1), O_0 = I in M_1."""
    tables, prev = [], np.repeat(np.eye(3)[None], len(E), axis=0)
    for slot, rho in enumerate(rhos):
        lam, V = spectrum(Hr, E, np.full(len(E), rho), antinu)
        rate = [(-LOE * (lam[:, k] - lam[:, 0]) / E) / np.pi for k in (1, 2)]
        v1, v2 = V[:, :, 1], V[:, :, 2]
        O = np.stack([np.cross(v1, v2), v1, v2], axis=2)          # columns v_0, v_1, v_2
        M = np.einsum("eak,eac->ekc", O, prev).astype(F32) if slot >= 1 else None
        tables.append((rate, [v1.astype(F32), v2.astype(F32)], M))
        if slot >= 1:
            prev = O
    return tables


def cross3(a, b):
    """Row-wise cross product of [n, 3] arrays."""
    return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1], a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                     a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], axis=1)


def local_basis_tables(tables):
    """The batched kernel's basis changes, formed in float32 from the stored float32 v_1, v_2:
    O_s = [v_1 x v_2, v_1, v_2], M_s = O_s^T O_{s-1}, M_1 = O_1^T."""
    out, prev = [], None
    for slot, (rate, v, _) in enumerate(tables):
        if slot == 0:
            out.append((rate, v, None))
            continue
        v1, v2 = v[0].astype(F32), v[1].astype(F32)
        O = np.stack([cross3(v1, v2).astype(F32), v1, v2], axis=2)          # [e][component][column]
        M = np.transpose(O, (0, 2, 1)).copy() if prev is None else np.einsum("eak,eac->ekc", O, prev).astype(F32)
        out.append((rate, v, M))
        prev = O
    return out


def emulate_eigenbasis(g, shuffle_m=False, det_phase=True, batched=False):
    """The single-grid mixed kernel: for rows with L >= 2 the first two columns of the running
    product Z = D_i M_i ... D_1 M_1 (M real 3x3 basis changes, D_i = diag(1, e^{i dphi_1},
    e^{i dphi_2})) and the FP64 sum of the phases (half turns), then the first two columns of
    Z' = M_L Z, normalised, completed by det(Z') conj(c_0 x c_1) with det = e^{i pi sum}, and
    T[x][y] = sum_c (Z' a_x)_c D_L,c (Z' G)_c,y with a_e = e_0, a_mu = row mu of W; for L = 0 the
    vacuum form T = I + sum_{k=1,2} w_k U_xk conj(U_yk); L = 1 as in emulate().  batched: the
    batched kernel's variant, basis changes formed from the float eigenvectors
    (local_basis_tables) and L = 0 as in emulate()."""
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
    U18 = np.asarray(p.Mix_U)
    Hr, W = real_basis((U18[0::2] + 1j * U18[1::2]).reshape(3, 3), np.asarray(p.dm))
    Wc = np.conj(W[:2, :].T).astype(np.complex64)
    Wr = W[:2, :].astype(np.complex64)
    rhos = [0.0] + [float(r) * 0.5 for r in p.rhos]
    tables = eigenbasis_tables(Hr, E, rhos, antinu)
    if batched:
        tables = local_basis_tables(tables)
    base = emulate(g)
    # vacuum constants (W v_k)_x conj((W v_k)_y) with the table's labelling of the eigenvectors (for
    # the kernel's labelling, by mass index, these are U_xk conj(U_yk))
    _, V0 = spectrum(Hr, E[:1], np.zeros(1), antinu)
    wv = [(W @ V0[0, :, k])[:2] for k in (1, 2)]
    vac = [np.outer(x, np.conj(x)).astype(np.complex64) for x in wv]
    out = np.zeros((len(p.cosineList), n_e, 9))
    for ic, cz in enumerate(p.cosineList):
        L = int(p.maxlayers[ic])
        dist = oracle.layer_distances(p.radii, L, cz, float(p.ProductionHeightinCentimeter)) / 1e5
        if L == 0 and not batched:       # one vacuum layer: T = I + sum_k w_k U_xk conj(U_yk)
            w = [phase_factor_m1(tables[0][0][k] * dist[0]) for k in (0, 1)]
            T = (np.eye(2, dtype=np.complex64)[None] + w[0][:, None, None] * vac[0][None] +
                 w[1][:, None, None] * vac[1][None]).astype(np.complex64)
            P2 = (np.abs(T) ** 2).astype(F32)
            P = np.zeros((n_e, 3, 3), F32)
            P[:, :2, :2] = np.transpose(P2, (0, 2, 1))
            P[:, 0, 2] = np.clip(1 - P[:, 0, 0] - P[:, 0, 1], 0, 1)
            P[:, 1, 2] = np.clip(1 - P[:, 1, 0] - P[:, 1, 1], 0, 1)
            P[:, 2, 0] = np.clip(1 - P[:, 0, 0] - P[:, 1, 0], 0, 1)
            P[:, 2, 1] = np.clip(1 - P[:, 0, 1] - P[:, 1, 1], 0, 1)
            P[:, 2, 2] = np.clip(1 - P[:, 2, 0] - P[:, 2, 1], 0, 1)
            out[ic] = P.reshape(n_e, 9)
            continue
        if L < 2:
            out[ic] = base[ic]
            continue
        tsum = np.zeros(n_e)

        def d(i, add=True):
            nonlocal tsum
            rate = tables[i][0]
            t1, t2 = rate[0] * dist[i], rate[1] * dist[i]
            if add:
                tsum = tsum + t1 + t2
            return [phase_factor(t1), phase_factor(t2)]

        def Mof(i):
            M = tables[i][2]
            return M[:, ::-1, :] if shuffle_m else M

        d1 = d(1)
        Z = Mof(1)[:, :, :2].astype(np.complex64)
        Z[:, 1, :] *= d1[0][:, None]
        Z[:, 2, :] *= d1[1][:, None]
        for i in range(2, L):
            Z = (Mof(i) @ Z).astype(np.complex64)
            di = d(i)
            Z[:, 1, :] *= di[0][:, None]
            Z[:, 2, :] *= di[1][:, None]
        Zp = (Mof(L) @ Z).astype(np.complex64)
        Zp = (Zp / np.sqrt((np.abs(Zp) ** 2).sum(1, keepdims=True)).astype(F32)).astype(np.complex64)
        det = phase_factor(tsum) if det_phase else np.ones(n_e, np.complex64)
        c2 = (det[:, None] * np.conj(cross3(Zp[:, :, 0], Zp[:, :, 1]))).astype(np.complex64)
        Zp = np.concatenate([Zp, c2[:, :, None]], axis=2)
        dL = d(L, add=False)
        G = apply_layer(np.repeat(Wc[None], n_e, axis=0), tables[0][1], [phase_factor_m1(tables[0][0][k] * dist[0])
                                                                           for k in (0, 1)])
        Ug = (Zp @ G).astype(np.complex64)                                   # [e][c][y]
        Ua = (Zp @ Wr.T[None]).astype(np.complex64)                          # [e][c][x]
        Ua[:, 1, :] *= dL[0][:, None]
        Ua[:, 2, :] *= dL[1][:, None]
        T = np.einsum("ecx,ecy->exy", Ua, Ug).astype(np.complex64)           # T[e, out, in]
        P2 = (np.abs(T) ** 2).astype(F32)
        P = np.zeros((n_e, 3, 3), F32)
        P[:, :2, :2] = np.transpose(P2, (0, 2, 1))
        P[:, 0, 2] = np.clip(1 - P[:, 0, 0] - P[:, 0, 1], 0, 1)
        P[:, 1, 2] = np.clip(1 - P[:, 1, 0] - P[:, 1, 1], 0, 1)
        P[:, 2, 0] = np.clip(1 - P[:, 0, 0] - P[:, 1, 0], 0, 1)
        P[:, 2, 1] = np.clip(1 - P[:, 0, 1] - P[:, 1, 1], 0, 1)
        P[:, 2, 2] = np.clip(1 - P[:, 2, 0] - P[:, 2, 1], 0, 1)
        out[ic] = P.reshape(n_e, 9)
    return out


@pytest.mark.parametrize("batched", [False, True])
@pytest.mark.parametrize("name", golden_names())
def test_eigenbasis_algebra_meets_tolerance(name, batched):
    """The eigenbasis form of the running product (rows with L >= 2), with the table's basis
    changes (single-grid kernels) or with basis changes formed from the float eigenvectors
    (batched kernel), against the reference's golden values, within the mixed bar."""
    g = load_golden(name)
    P = emulate_eigenbasis(g, batched=batched)
    err = np.abs(P - g["P"]).max()
    print(f"{name} batched={batched}: emulated eigenbasis algebra max|dP| {err:.2e}")
    assert err <= MIXED_TOL


def test_eigenbasis_emulation_detects_a_wrong_basis_change():
    """M_s = O_s^T O_{s-1}: using its rows in the wrong order breaks the bar."""
    g = load_golden("prem59_nu")
    bad = emulate_eigenbasis(g, shuffle_m=True)
    assert np.abs(bad - g["P"]).max() > 1e-2


def test_eigenbasis_emulation_detects_a_missing_determinant():
    """The third column of Z' is det(Z') conj(c_0 x c_1): leaving out the determinant's phase
    breaks the bar."""
    g = load_golden("prem59_nu")
    bad = emulate_eigenbasis(g, det_phase=False)
    assert np.abs(bad - g["P"]).max() > 1e-2


def test_eigenbasis_long_paths_on_a_100_shell_earth():
    """Paths of up to 199 layers (100 equal shells): the completed two-column product stays within
    the bar against the oracle (the normalisation of the columns removes the norm drift that
    dominates over so many products)."""
    from cudaprob3_amd import workloads as W
    radii, rhos = W.fine100_model()
    cos = np.linspace(-1.0, -0.3, 12)
    E = np.geomspace(0.1, 1000.0, 48)
    for antinu, dm32 in ((0, 2.5e-3), (1, -2.5e-3)):
        params = list(W.NUFIT)
        params[5] = dm32
        g = dict(cosines=cos, energies=E, radii=radii[::-1].copy(), rhos=rhos[::-1].copy(), h_km=22.0,
                 params=np.array(params), antinu=antinu)
        g["P"] = oracle.propagate(antinu, cos, E, g["radii"], g["rhos"], 22.0, params)
        for batched in (False, True):
            err = np.abs(emulate_eigenbasis(g, batched=batched) - g["P"]).max()
            print(f"fine100 antinu={antinu} dm32={dm32} batched={batched}: emulated max|dP| {err:.2e}")
            assert err <= 0.5 * MIXED_TOL
