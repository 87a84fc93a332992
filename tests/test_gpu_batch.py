"""Batched parameter scans (one persistent launch over many sets with uneven grids and Earth
models) against the single-grid mixed propagator (bit for bit: same per-pair code and
tables) and the oracle (mixed tolerance)."""
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
            np.testing.assert_array_equal(got, single(s))
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
            np.testing.assert_array_equal(got, single(s))
            rows = np.arange(0, len(s.cosines), 23)
            ref = oracle_mod.propagate(s.antinu, s.cosines[rows], s.energies, s.radii, s.rhos, s.h_km, s.params)
            assert np.abs(got[:, rows, :].transpose(1, 2, 0) - ref).max() <= MIXED_TOL
