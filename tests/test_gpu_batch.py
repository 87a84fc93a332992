"""Batched parameter scans (one persistent launch over many sets with uneven grids and Earth
models) against the single-grid mixed propagator and the oracle (mixed tolerance).  The two
kernels share the per-pair code of rows that cross exactly one shell (bit for bit there); for
deeper rows both carry the running product in the shells' eigenbases, with basis changes from
the table (single grid, FP64-formed) or formed from the float eigenvectors (batched), and for
rows that cross no shell the single-grid kernel uses the vacuum form, so there they agree within
the mixed bar."""
import numpy as np
import pytest

from cudaprob3_amd import workloads as W
from cudaprob3_amd.batch import BatchPropagator, SetSpec
from cudaprob3_amd.propagator import HipPropagatorSingle

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-5


def small_sets(n=12, seed=21):
    rng = np.random.Generator(np.random.PCG64(seed))
    sets = []
    for i in range(n):
        p = W.random_params(rng)
        n_c = int(rng.integers(1, 40))
        n_e = int(rng.integers(1, 150))
        radii, rhos = W.random_earth(rng, int(rng.integers(4, 51)))
        if i % 5 == 3:           # a model without an r = 0 row
            radii, rhos = radii[2:], rhos[2:]
        sets.append(SetSpec(np.linspace(-1, 1, n_c) if n_c > 1 else np.array([-0.6]),
                            np.geomspace(0.1, 100, n_e) if n_e > 1 else np.array([3.0]),
                            radii, rhos, p, antinu=i % 2))
    return sets


def same_as_single(got, spec):
    """got (a set's batched results) against the single-grid propagator: bitwise on rows crossing
    exactly one shell, within the mixed bar on the others."""
    from cudaprob3_amd import metrics as M
    ref = single(spec)
    one = M.maxlayers_of(spec.cosines, spec.radii, spec.rhos) == 1
    np.testing.assert_array_equal(got[:, one, :], ref[:, one, :])
    assert np.abs(got - ref).max() <= MIXED_TOL


def single(spec):
    with HipPropagatorSingle(0, len(spec.cosines), len(spec.energies), precision="mixed") as p:
        p.setEnergyList(spec.energies)
        p.setCosineList(spec.cosines)
        p.setMNSMatrix(*spec.params[:4])
        p.setNeutrinoMasses(*spec.params[4:])
        p.setDensity(spec.radii, spec.rhos)
        p.setProductionHeight(spec.h_km)
        p.calculateProbabilities(spec.antinu)
        return p.getProbabilities().copy()


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_single_and_oracle(oracle_mod, graph):
    sets = small_sets()
    print(f"[gpu] batch of {len(sets)} sets, graph={graph}", flush=True)
    with BatchPropagator(sets, graph=graph) as b:
        b.compute()
        for i, s in enumerate(sets):
            got = b.results(i)
            same_as_single(got, s)
            ref = oracle_mod.propagate(s.antinu, s.cosines, s.energies, s.radii, s.rhos, s.h_km, s.params)
            assert np.abs(got.transpose(1, 2, 0) - ref).max() <= MIXED_TOL


def test_graph_replays_follow_parameter_changes(oracle_mod):
    """Scan steps: new hypotheses written between replays; every replay equals the eager
    (no-graph) computation of the same hypotheses."""
    sets = small_sets(8, seed=5)
    rng = np.random.Generator(np.random.PCG64(77))
    with BatchPropagator(sets, graph=True) as g, BatchPropagator(sets, graph=False) as e:
        for step in range(4):
            print(f"[gpu] graph replay step {step}", flush=True)
            for i in range(len(sets)):
                p = W.random_params(rng)
                g.set_params(i, p)
                e.set_params(i, p)
            g.compute()
            e.compute()
            for i in range(len(sets)):
                np.testing.assert_array_equal(g.results(i), e.results(i))
            s = g.sets[step]
            ref = oracle_mod.propagate(s.antinu, s.cosines, s.energies, s.radii, s.rhos, s.h_km, s.params)
            assert np.abs(g.results(step).transpose(1, 2, 0) - ref).max() <= MIXED_TOL


def test_autotune_and_configs_agree():
    sets = small_sets(10, seed=9)
    with BatchPropagator(sets, graph=False) as b:
        w, gpc, ms = b.autotune()
        assert w in (4, 8, 16) and 1 <= gpc <= 8 and ms > 0
        b.compute()
        base = [b.results(i) for i in range(len(sets))]
        for cfg in [(4, 1), (16, 2), (8, 8)]:
            b.configure(*cfg)
            b.compute()
            for i in range(len(sets)):
                np.testing.assert_array_equal(b.results(i), base[i])


def test_scan1024_full_size(oracle_mod):
    """The 1024-set scan (64-512 cosines x 64-1024 energies, 4-50 shells each) in one launch;
    eight sets checked against the single propagator and the oracle on strided rows."""
    specs = [SetSpec(s.cosines, s.energies, s.radii, s.rhos, s.params, s.antinu, s.h_km) for s in W.scan_sets()]
    print(f"[gpu] scan1024: {sum(len(s.cosines) * len(s.energies) for s in specs)} pairs", flush=True)
    with BatchPropagator(specs, graph=True) as b:
        b.compute()
        b.compute()
        for i in range(0, 1024, 131):
            s = specs[i]
            got = b.results(i)
            same_as_single(got, s)
            rows = np.arange(0, len(s.cosines), 23)
            ref = oracle_mod.propagate(s.antinu, s.cosines[rows], s.energies, s.radii, s.rhos, s.h_km, s.params)
            assert np.abs(got[:, rows, :].transpose(1, 2, 0) - ref).max() <= MIXED_TOL


@pytest.mark.parametrize("graph", [False, True])
def test_binned_mode_matches_unbinned_average(oracle_mod, graph):
    """Binned mode (bin averages reduced across lanes in the pair kernel, written into disjoint
    blocks of one strided array) against the unbinned results averaged afterwards and against
    the oracle's FP64 average; cells outside the blocks stay untouched."""
    import torch
    from cudaprob3_amd.surrogate import gl_energies
    rng = np.random.Generator(np.random.PCG64(8))
    radii, rhos = W.jitter12_model()
    sets, shapes = [], []
    for i in range(6):
        n_c, n_b = int(rng.integers(1, 9)), int(rng.integers(1, 11))
        cos = np.sort(rng.uniform(-1, 1, n_c))
        centers = np.sort(10 ** rng.uniform(-1.3, 3, n_b))
        E, w = gl_energies(centers)
        sets.append(SetSpec(cos, E, radii, rhos, W.random_params(rng), antinu=i % 2))
        shapes.append((n_c, n_b))
    print(f"[gpu] binned batch mode, graph={graph}", flush=True)
    # set i fills rows [r0_i, r0_i + n_c) x bins [0, n_b) of a [9][R][B] array with B = 16
    R, B = sum(s[0] for s in shapes) + 3, 16
    out = torch.full((9, R, B), -7.0, dtype=torch.float32, device="cuda")
    r0 = np.cumsum([0] + [s[0] for s in shapes])[:-1] + 1
    with BatchPropagator(sets, graph=False) as b:
        b.compute()
        plain = [b.results(i) for i in range(len(sets))]
    with BatchPropagator(sets, graph=graph) as b:
        torch.cuda.synchronize()
        b.set_binning(16, w, out.data_ptr(), r0 * B, R * B, B)
        b.compute()
        b.compute()                         # a replay writes the same values
        got = out.cpu().numpy()
    mask = np.zeros((R, B), bool)
    for i, (s, (n_c, n_b)) in enumerate(zip(sets, shapes)):
        blk = got[:, r0[i]:r0[i] + n_c, :n_b]
        avg = np.einsum("tcbg,g->tcb", plain[i].reshape(9, n_c, n_b, 16).astype(np.float64), w.astype(np.float64))
        assert np.abs(blk - avg).max() <= 2e-6
        ref = oracle_mod.propagate(s.antinu, s.cosines, s.energies, radii, rhos, 22.0, s.params)
        ref = np.einsum("cbgt,g->tcb", ref.reshape(n_c, n_b, 16, 9), w.astype(np.float64))
        assert np.abs(blk - ref).max() <= MIXED_TOL
        mask[r0[i]:r0[i] + n_c, :n_b] = True
    assert np.all(got[:, ~mask] == -7.0)


def test_binned_mode_rejects_bad_layouts():
    from cudaprob3_amd._lib import Cp3hError
    from cudaprob3_amd.surrogate import gl_energies
    radii, rhos = W.jitter12_model()
    E, w = gl_energies([1.0, 2.0])
    with BatchPropagator([SetSpec(np.array([-0.5, 0.5]), E, radii, rhos, W.NUFIT)], graph=False) as b:
        import torch
        out = torch.zeros(64, device="cuda")
        with pytest.raises(Cp3hError):
            b.set_binning(3, w[:3], out.data_ptr(), [0], 4, 2)          # not a power of two
        b.set_binning(32, np.ones(32) / 32, out.data_ptr(), [0], 2, 1)   # 32 energies = one bin per row
        with pytest.raises(Cp3hError):
            b.set_binning(16, w, out.data_ptr(), [0], 4, 1)            # row stride below the bin count
        with pytest.raises(Cp3hError):
            b.set_binning(16, w, out.data_ptr(), [0], 2, 2)            # planes overlap
        b.set_binning(0, None, 0, 0, 0, 0)                              # back to unbinned results
        b.compute()
        assert b.results(0).shape == (9, 2, 32)


@pytest.mark.parametrize("graph", [False, True])
def test_row_limits(graph):
    """Per-row energy limits: each row's leading energies (bins in binned mode) equal the
    unlimited computation bit for bit; nothing past a row's limit is written."""
    import torch
    from cudaprob3_amd.surrogate import gl_energies
    rng = np.random.Generator(np.random.PCG64(12))
    radii, rhos = W.jitter12_model()
    cos = np.linspace(-0.98, 0.97, 37)
    centers = np.geomspace(0.07, 700, 21)
    E, w = gl_energies(centers)
    sets = [SetSpec(cos, E, radii, rhos, W.random_params(rng), antinu=i) for i in range(2)]
    lim_bins = rng.integers(0, 22, (2, len(cos))).astype(np.int32)
    lim_bins[0, 3] = 0
    lim_bins[1, 5] = 21
    print(f"[gpu] batch row limits, graph={graph}", flush=True)
    with BatchPropagator(sets, graph=False) as b:
        b.compute()
        full = [b.results(i) for i in range(2)]
    nb = len(centers)
    out = torch.full((2, 9, len(cos), nb), -5.0, device="cuda")
    with BatchPropagator(sets, graph=graph) as b:
        for i in range(2):
            b.set_row_limits(i, 16 * lim_bins[i])
        torch.cuda.synchronize()
        b.set_binning(16, w, out.data_ptr(), [0, 9 * len(cos) * nb], len(cos) * nb, nb)
        b.compute()
        b.compute()
        got = out.cpu().numpy()
        b.set_binning(0, None, 0, 0, 0, 0)
        b.compute()
        limited = [b.results(i) for i in range(2)]
    for i in range(2):
        keep = np.arange(nb)[None, :] < lim_bins[i][:, None]                        # [ic][bin]
        avg = np.einsum("tcbg,g->tcb", full[i].reshape(9, len(cos), nb, 16).astype(np.float64), w.astype(np.float64))
        assert np.abs(got[i][:, keep] - avg[:, keep]).max() <= 2e-6
        assert np.all(got[i][:, ~keep] == -5.0)
        keep_e = np.repeat(keep, 16, axis=1)
        np.testing.assert_array_equal(limited[i][:, keep_e], full[i][:, keep_e])


def test_device_results_layout():
    """Device results: every cosine row starts a 128-byte line (row stride = n_e rounded up to
    32 floats); read through the reported strides they equal the dense copy."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    sets = small_sets(6, seed=31)
    print(f"[gpu] device results layout, {len(sets)} sets", flush=True)
    with BatchPropagator(sets, graph=False) as b:
        b.compute()
        for i, s in enumerate(sets):
            plane, row = b.result_strides(i)
            n_c, n_e = len(s.cosines), len(s.energies)
            assert row == -(-n_e // 32) * 32 and plane == n_c * row
            ptr = b.device_results(i)
            assert ptr % 128 == 0
            dense = b.results(i)
            raw = np.empty(9 * plane, np.float32)
            assert hip.hipMemcpy(raw.ctypes.data, ptr, raw.nbytes, 2) == 0      # device to host
            np.testing.assert_array_equal(raw.reshape(9, n_c, row)[:, :, :n_e], dense)
