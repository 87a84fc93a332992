"""Parity of the MI355X kernels (through libcp3hip.so) with the oracle and the reference's
golden vectors.

Tolerances (max |dP| over all nine channels):
  * fp64_plain, fp64 : FP64_TOL = 1e-12 against the reference / oracle FP64 CPU path, with one
                       named exception: on the golden fixture degenerate_dm23_zero (dm23 = 0,
                       whose 5e-9 eV^2 degeneracy nudge, propagator.hpp:233-236, amplifies libm
                       rounding) the bar is twice the reference algorithm's own drift under a
                       one-ulp change of libm (oracle.libm_sensitivity), below 1e-11;
  * mixed            : MIXED_TOL = 1e-5 against the same FP64 results (FP32 bulk arithmetic,
                       FP64 eigenvalues and phase reduction).
"""
import numpy as np
import pytest

from conftest import golden_names, load_golden
from cudaprob3_amd import workloads as W
from cudaprob3_amd.propagator import HipPropagatorSingle, Neutrino, Antineutrino

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-5


# the one input whose FP64 bar follows the reference's libm sensitivity instead of 1e-12
LIBM_RELAXED = {"degenerate_dm23_zero"}


def tolerance(precision, oracle_mod=None, reference=None, *inputs, name=None):
    if precision == "mixed":
        return MIXED_TOL
    if name in LIBM_RELAXED:
        tol = oracle_mod.fp64_tolerance(oracle_mod.libm_sensitivity(*inputs, reference=reference))
        assert tol < 1e-11
        return tol
    return FP64_TOL


def run_gpu(precision, cosines, energies, radii, rhos, params, h_km, antinu, device=0):
    dtype = np.float64
    with HipPropagatorSingle(device, len(cosines), len(energies), dtype=dtype, precision=precision) as p:
        p.setEnergyList(energies)
        p.setCosineList(cosines)
        p.setMNSMatrix(*params[:4])
        p.setNeutrinoMasses(*params[4:])
        p.setDensity(radii, rhos)
        p.setProductionHeight(h_km)
        p.calculateProbabilities(Antineutrino if antinu else Neutrino)
        return np.ascontiguousarray(p.getProbabilities().transpose(1, 2, 0))   # -> [ic, ie, t]


@pytest.mark.parametrize("precision", ["fp64_plain", "fp64", "mixed"])
@pytest.mark.parametrize("name", golden_names())
def test_golden(oracle_mod, precision, name):
    g = load_golden(name)
    print(f"[gpu] golden {name} {precision}", flush=True)
    P = run_gpu(precision, g["cosines"], g["energies"], g["radii"], g["rhos"], g["params"],
                float(g["h_km"]), int(g["antinu"]))
    err = np.abs(P - g["P"]).max()
    tol = tolerance(precision, oracle_mod, g["P"], int(g["antinu"]), g["cosines"], g["energies"], g["radii"],
                    g["rhos"], float(g["h_km"]), g["params"], name=name)
    print(f"[gpu]   max|dP| {err:.3e} (bar {tol:.1e})")
    assert err <= tol, f"{name} {precision}: max|dP| = {err:.3e}"


@pytest.mark.parametrize("antinu", [0, 1])
@pytest.mark.parametrize("precision", ["fp64_plain", "fp64", "mixed"])
def test_prem4_256_full_grid(oracle_mod, precision, antinu):
    """256 cosZ x 256 energies, PREM 4 layers: every pair against the oracle."""
    g = W.prem4_256()
    print(f"[gpu] prem4_256 {precision} antinu={antinu}", flush=True)
    ref = oracle_mod.propagate(antinu, g.cosines, g.energies, g.radii, g.rhos, g.h_km, g.params)
    P = run_gpu(precision, g.cosines, g.energies, g.radii, g.rhos, g.params, g.h_km, antinu)
    err = np.abs(P - ref)
    tol = tolerance(precision)
    assert tol == (MIXED_TOL if precision == "mixed" else 1e-12)
    print(f"[gpu] prem4_256 {precision} antinu={antinu}: max {err.max():.3e} rms {np.sqrt((err**2).mean()):.3e}"
          f" (bar {tol:.1e})")
    assert err.max() <= tol


@pytest.mark.parametrize("n_c,n_e", [(1, 1), (3, 37), (5, 64), (7, 65), (2, 130), (0, 5), (4, 0)])
@pytest.mark.parametrize("precision", ["fp64_plain", "fp64", "mixed"])
def test_ragged_and_empty_grids(oracle_mod, precision, n_c, n_e):
    cos = np.linspace(-1, 1, n_c) if n_c > 1 else np.full(n_c, -0.7)
    E = np.geomspace(0.2, 40, n_e) if n_e > 1 else np.full(n_e, 2.0)
    print(f"[gpu] ragged {n_c}x{n_e} {precision}", flush=True)
    P = run_gpu(precision, cos, E, W.PREM12[:, 0], W.PREM12[:, 1], W.NUFIT, 22.0, 0)
    assert P.shape == (n_c, n_e, 9)
    if n_c and n_e:
        ref = oracle_mod.propagate(0, cos, E, W.PREM12[:, 0], W.PREM12[:, 1], 22.0, W.NUFIT)
        assert np.abs(P - ref).max() <= tolerance(precision)


def test_mixed_rows_together_equal_rows_alone():
    """The mixed kernel propagates two cosine rows per wavefront together when they cross the
    same number of shells (sharing table reads) and one after the other otherwise; every row of a
    grid must equal, bit for bit, the same row propagated alone (a one-row grid).  The rows come
    in pairs (a wavefront's two rows): for every shell count L = 0..12 two rows inside the same
    band, then pairs straddling each band edge, then an odd last row."""
    radii = W.PREM12[:, 0]
    R = radii.max()
    inner = np.sort(radii[(radii > 0) & (radii < R)])[::-1]
    edges = np.concatenate([[0.0], -np.sqrt(1 - (inner / R) ** 2), [-1.0]])   # L = k for edges[k+1] <= cos < edges[k]
    same = [x for k in range(len(edges) - 1) for x in np.linspace(edges[k + 1], edges[k], 4)[1:3]]
    straddle = [x for e in edges[1:-1] for x in (e - 1e-4, e + 1e-4)]
    cos = np.array([0.5, 0.7] + same + straddle + [-0.33])
    E = np.geomspace(0.1, 100, 70)
    print(f"[gpu] mixed rows together vs alone: {len(cos)} rows x {len(E)}", flush=True)
    full = run_gpu("mixed", cos, E, W.PREM12[:, 0], W.PREM12[:, 1], W.NUFIT, 22.0, 0)
    for i, c in enumerate(cos):
        alone = run_gpu("mixed", np.array([c]), E, W.PREM12[:, 0], W.PREM12[:, 1], W.NUFIT, 22.0, 0)
        np.testing.assert_array_equal(full[i], alone[0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_jitter12_2048_full_size(oracle_mod, precision):
    """The 2048 x 2048 jittered 12-shell workload, all four nu/nubar x NO/IO variants: strided rows
    against the oracle; for mixed, every one of the 4.2 M pairs per variant against the FP64
    kernel (held to the oracle at 1e-12 by the fp64 case), and the directly computed entries
    (P_ee, P_emu, P_mue, P_mumu; the other five follow from unitarity) summing to at most one
    along each row and column; for fp64, whose nine entries are all computed, unitarity."""
    for g in W.jitter12_variants(2048):
        print(f"[gpu] jitter12_2048 {precision} antinu={g.antinu} dm32={g.params[5]}", flush=True)
        P = run_gpu(precision, g.cosines, g.energies, g.radii, g.rhos, g.params, g.h_km, g.antinu)
        rows = np.arange(0, 2048, 37)
        ref = oracle_mod.propagate(g.antinu, g.cosines[rows], g.energies, g.radii, g.rhos, g.h_km, g.params)
        err = np.abs(P[rows] - ref)
        tol = tolerance(precision)
        print(f"[gpu]   vs oracle on {len(rows)} rows: max {err.max():.3e} rms {np.sqrt((err**2).mean()):.3e} "
              f"(bar {tol:.1e})")
        assert err.max() <= tol
        M = P.reshape(2048, 2048, 3, 3)
        if precision == "mixed":
            P64 = run_gpu("fp64", g.cosines, g.energies, g.radii, g.rhos, g.params, g.h_km, g.antinu)
            e64 = np.abs(P - P64)
            print(f"[gpu]   vs fp64 kernel, all pairs: max {e64.max():.3e} rms {np.sqrt((e64 ** 2).mean()):.3e}")
            assert e64.max() <= MIXED_TOL
            del P64, e64
            direct = M[:, :, :2, :2]
            over = max((direct.sum(axis=3) - 1).max(), (direct.sum(axis=2) - 1).max())
            assert over <= MIXED_TOL
        else:
            uni = max(np.abs(M.sum(axis=3) - 1).max(), np.abs(M.sum(axis=2) - 1).max())
            assert uni <= 1e-11


@pytest.mark.parametrize("antinu,inverted", [(0, False), (1, False), (0, True)])
def test_fine100_mixed_long_paths(oracle_mod, antinu, inverted):
    """The 100-shell Earth of the Markov-chain workload, whose core-crossing paths have up to 199
    layers (the longest running products of any workload): the mixed kernel on a 512 x 512 grid
    over the workload's ranges, all rows against the FP64 kernel and strided rows against the
    oracle, within the mixed bar."""
    g = W.fine100_grid(512)
    params = list(g.params)
    if inverted:
        params[5] = -params[5]
    print(f"[gpu] fine100 512^2 mixed antinu={antinu} dm32={params[5]}", flush=True)
    P = run_gpu("mixed", g.cosines, g.energies, g.radii, g.rhos, params, g.h_km, antinu)
    P64 = run_gpu("fp64", g.cosines, g.energies, g.radii, g.rhos, params, g.h_km, antinu)
    rows = np.arange(0, 512, 29)
    ref = oracle_mod.propagate(antinu, g.cosines[rows], g.energies, g.radii, g.rhos, g.h_km, params)
    e64 = np.abs(P - P64)
    eref = np.abs(P[rows] - ref).max()
    print(f"[gpu]   vs fp64 kernel max {e64.max():.3e} rms {np.sqrt((e64 ** 2).mean()):.3e}; vs oracle {eref:.3e}")
    assert e64.max() <= MIXED_TOL and eref <= MIXED_TOL


def test_mixed_extreme_phases():
    """Energies down to 1e-8 GeV on Earth-crossing paths give phase differences of up to ~10^10 rad
    (more than 2^31 half turns): the mixed kernel's sign (-1)^q must still follow the parity of
    q.  Energies up to 1e6 GeV give tiny phases and matter-dominated spectra.  Against the FP64
    kernel (same FP64 phase range) within the mixed bar."""
    cos = np.array([-1.0, -0.83, -0.5, -0.1, 0.3])
    E = np.array([1e-8, 2.3e-8, 1e-7, 1e-6, 1e-4, 1e-2, 1.0, 1e3, 1e4, 1e5, 1e6])
    print("[gpu] mixed, extreme phases", flush=True)
    for antinu in (0, 1):
        P = run_gpu("mixed", cos, E, W.PREM12[:, 0], W.PREM12[:, 1], W.NUFIT, 22.0, antinu)
        P64 = run_gpu("fp64", cos, E, W.PREM12[:, 0], W.PREM12[:, 1], W.NUFIT, 22.0, antinu)
        err = np.abs(P - P64).max()
        print(f"[gpu]   antinu {antinu}: max|dP| vs fp64 kernel {err:.3e}")
        assert err <= MIXED_TOL


def test_api_errors_and_state():
    with pytest.raises(RuntimeError, match="not available"):
        HipPropagatorSingle(999, 4, 4)
    p = HipPropagatorSingle(4, 4)
    p.setEnergyList(np.geomspace(1, 10, 4))
    p.setCosineList(np.linspace(-1, 0, 4))
    p.setMNSMatrix(*W.NUFIT[:4])
    p.setNeutrinoMasses(*W.NUFIT[4:])
    with pytest.raises(RuntimeError, match="production height was not set"):
        p.calculateProbabilities(Neutrino)
    p.setProductionHeight(22.0)
    with pytest.raises(RuntimeError, match="density was not set"):
        p.calculateProbabilities(Neutrino)
    p.setDensity(W.PREM4[:, 0], W.PREM4[:, 1])
    p.calculateProbabilities(Neutrino)
    with pytest.raises(RuntimeError, match="Invalid indices"):
        p.getProbability(4, 0, 0)
    v = p.getProbability(3, 3, 8)
    assert 0.0 <= v <= 1.0
    with pytest.raises(RuntimeError, match="timing enabled"):
        p.last_compute_ms()                 # timing is opt-in
    p.set_timing(True)
    p.calculateProbabilities(Neutrino)
    assert p.last_compute_ms() > 0.0
    p.close()
    with pytest.raises(RuntimeError, match="moved from"):
        p.calculateProbabilities(Neutrino)


def test_independent_handles_do_not_share_parameters(oracle_mod):
    """Two propagators with different hypotheses, launched back to back without waiting,
    keep their own parameters (the reference's process-global constants would race)."""
    g = W.prem4_256()
    cos, E = g.cosines[::4], g.energies[::4]
    hyp = [W.NUFIT, (0.6, 0.14, 0.9, 0.3, 7.0e-5, -2.4e-3)]
    props = []
    for h in hyp:
        p = HipPropagatorSingle(len(cos), len(E), precision="fp64")
        p.setEnergyList(E); p.setCosineList(cos); p.setMNSMatrix(*h[:4]); p.setNeutrinoMasses(*h[4:])
        p.setDensity(g.radii, g.rhos); p.setProductionHeight(22.0)
        props.append(p)
    for p in props:
        p.calculateProbabilitiesAsync(Neutrino)
    for p, h in zip(props, hyp):
        p.waitForCompletion()
        ref = oracle_mod.propagate(0, cos, E, g.radii, g.rhos, 22.0, h)
        assert np.abs(p.getProbabilities().transpose(1, 2, 0) - ref).max() <= tolerance("fp64")
        p.close()


def test_concurrent_streams_equal_one_at_a_time():
    """The benchmark's step: the four jittered-Earth variants (nu/nubar x NO/IO) on their own
    propagators, each on its own stream, all launched before any is waited for, repeated a few
    times, so that one variant's set-up kernel runs beside another's pair kernel.  Every result
    equals, bit for bit, the same propagator's result computed alone."""
    grids = W.jitter12_variants(512)
    props = []
    for g in grids:
        p = HipPropagatorSingle(len(g.cosines), len(g.energies), precision="mixed")
        p.setEnergyList(g.energies); p.setCosineList(g.cosines); p.setDensity(g.radii, g.rhos)
        p.setProductionHeight(g.h_km); p.setMNSMatrix(*g.params[:4]); p.setNeutrinoMasses(*g.params[4:])
        props.append(p)
    alone = []
    for p, g in zip(props, grids):
        p.calculateProbabilities(g.antinu)
        alone.append(p.getProbabilities().copy())
    print("[gpu] four variants on four streams, 3 rounds", flush=True)
    for _ in range(3):
        for p, g in zip(props, grids):
            p.calculateProbabilitiesAsync(g.antinu)
        for p in props:
            p.waitForCompletion()
        for p, ref in zip(props, alone):
            np.testing.assert_array_equal(p.getProbabilities(), ref)
    for p in props:
        p.close()


@pytest.mark.parametrize("name", ["prem12_io_dcp", "prem12_io_dcp_nubar", "prem59_nu"])
def test_mixed_any_phase_convention(name):
    """The mixed kernels' real basis is derived from U itself (host_util.hpp real_basis), so a U
    with arbitrary row and column phases -- the same physics -- gives the same probabilities:
    within the mixed bar of the golden P for the PDG U."""
    g = load_golden(name)
    rng = np.random.default_rng(11)
    cos, E = g["cosines"], g["energies"]
    with HipPropagatorSingle(0, len(cos), len(E), precision="mixed") as p:
        p.setEnergyList(E); p.setCosineList(cos)
        p.setDensity(g["radii"], g["rhos"]); p.setProductionHeight(float(g["h_km"]))
        p.setMNSMatrix(*g["params"][:4]); p.setNeutrinoMasses(*g["params"][4:])
        U18 = np.asarray(p.Mix_U, dtype=np.float64)
        U = (U18[0::2] + 1j * U18[1::2]).reshape(3, 3)
        V = np.diag(np.exp(1j * rng.uniform(0, 6.3, 3))) @ U @ np.diag(np.exp(1j * rng.uniform(0, 6.3, 3)))
        assert np.abs(V - U).max() > 0.1
        p.Mix_U = list(np.stack([V.real, V.imag], -1).reshape(18))
        print(f"[gpu] rephased U, {name}", flush=True)
        p.calculateProbabilities(Antineutrino if int(g["antinu"]) else Neutrino)
        err = np.abs(p.getProbabilities().transpose(1, 2, 0) - g["P"]).max()
    print(f"[gpu]   max|dP| {err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("params", [(0.5836, 0.0, 0.7222, 3.86, 7.42e-5, 2.494e-3),
                                    (0.0, 0.0, 0.7222, 3.86, 7.42e-5, -2.494e-3),
                                    (0.5836, 0.1496, 0.0, 0.0, 7.42e-5, 2.494e-3)])
def test_mixed_real_basis_edge_angles(oracle_mod, params):
    """Mixing angles at which the real basis has a free phase (theta13 = 0: (W^dagger M W)_12 = 0)
    or no first-column direction (theta12 = theta13 = 0: M_10 = M_20 = 0), and theta23 = 0:
    the mixed kernel against the oracle."""
    g = W.prem4_256()
    cos, E = g.cosines[::8], g.energies[::8]
    print(f"[gpu] edge angles {params}", flush=True)
    for antinu in (0, 1):
        with HipPropagatorSingle(0, len(cos), len(E), precision="mixed") as p:
            p.setEnergyList(E); p.setCosineList(cos)
            p.setMNSMatrix(*params[:4]); p.setNeutrinoMasses(*params[4:])
            p.setDensity(g.radii, g.rhos); p.setProductionHeight(22.0)
            p.calculateProbabilities(Antineutrino if antinu else Neutrino)
            P = p.getProbabilities().transpose(1, 2, 0)
        ref = oracle_mod.propagate(antinu, cos, E, g.radii, g.rhos, 22.0, params)
        err = np.abs(P - ref).max()
        print(f"[gpu]   antinu {antinu}: max|dP| {err:.3e}")
        assert err <= 1e-5


def test_float32_api_runs_mixed(oracle_mod):
    g = W.prem4_256()
    cos, E = g.cosines[::8].astype(np.float32), g.energies[::8].astype(np.float32)
    with HipPropagatorSingle(0, len(cos), len(E), dtype=np.float32) as p:
        assert p.precision == "mixed"
        p.setEnergyList(E); p.setCosineList(cos)
        p.setMNSMatrix(*W.NUFIT[:4]); p.setNeutrinoMasses(*W.NUFIT[4:])
        p.setDensity(g.radii, g.rhos); p.setProductionHeight(22.0)
        p.calculateProbabilities(Neutrino)
        P = p.getProbabilities().transpose(1, 2, 0)
        assert P.dtype == np.float32
        # the FP64 oracle on the same float32 inputs, with float32-rounded parameters
        params = [float(np.float32(x)) for x in W.NUFIT]
        ref = oracle_mod.propagate(0, cos.astype(np.float64), E.astype(np.float64), g.radii.astype(np.float32),
                                   g.rhos.astype(np.float32), 22.0, params)
        assert np.abs(P - ref).max() <= 5e-4   # float32 inputs and float32 host set-up dominate


def test_measured_fma_ceiling_is_plausible():
    """cp3h_measure_valu_peak: the sustained non-packed FMA rates lie below the data sheet's
    FP32 (157.3, packed) and FP64 (78.6) vector peaks and above a fifth of the non-packed rates, and
    the pair kernel's figure in bench.py is a fraction of them."""
    import ctypes
    from cudaprob3_amd import _lib
    lib = _lib.load()
    f32, f64 = ctypes.c_float(), ctypes.c_float()
    assert lib.cp3h_measure_valu_peak(0, ctypes.byref(f32), ctypes.byref(f64)) == 0, lib.cp3h_last_error()
    print(f"[gpu] measured FMA ceiling: FP32 {f32.value:.1f} TFLOP/s, FP64 {f64.value:.1f} TFLOP/s")
    assert 78.6 / 5 < f32.value < 157.3
    assert 78.6 / 5 < f64.value < 78.6 * 1.05
    assert lib.cp3h_measure_valu_peak(99, ctypes.byref(f32), None) == 3     # CP3H_ERR_DEVICE
