"""Host-side set-up of the Python Propagator mirror (cudaprob3_amd/propagator.py) against the
oracle's restatement of propagator.hpp, bit for bit, and the reference's error behaviour.
CPU only: the base class does no device work."""
import os

import numpy as np
import pytest

from cudaprob3_amd import workloads as W
from cudaprob3_amd.propagator import Propagator


def test_mns_and_masses_bit_exact(oracle_mod):
    rng = np.random.Generator(np.random.PCG64(1))
    for _ in range(50):
        p = W.random_params(rng)
        prop = Propagator(3, 3)
        prop.setMNSMatrix(*p[:4])
        prop.setNeutrinoMasses(*p[4:])
        np.testing.assert_array_equal(prop.Mix_U, oracle_mod.mns_matrix(*p[:4]))
        np.testing.assert_array_equal(prop.dm, oracle_mod.mass_matrix(*p[4:]))
    for dm in [(0.0, 2.5e-3), (7.4e-5, 0.0), (0.0, 0.0)]:    # degeneracy nudges
        prop = Propagator(1, 1)
        prop.setNeutrinoMasses(*dm)
        np.testing.assert_array_equal(prop.dm, oracle_mod.mass_matrix(*dm))


@pytest.mark.parametrize("model", ["PREM4", "PREM12", "PREM59", "desc", "norow0"])
def test_density_and_maxlayers_bit_exact(oracle_mod, model):
    if model == "desc":
        radii, rhos = W.PREM4[::-1, 0].copy(), W.PREM4[::-1, 1].copy()
    elif model == "norow0":
        radii, rhos = W.PREM12[3:, 0].copy(), W.PREM12[3:, 1].copy()
    else:
        tab = getattr(W, model)
        radii, rhos = tab[:, 0].copy(), tab[:, 1].copy()
    cos = np.concatenate([np.linspace(-1, 1, 301), [-0.0, 0.0, -1e-300]])
    prop = Propagator(len(cos), 4)
    prop.setCosineList(cos)
    prop.setDensity(radii, rhos)
    rd, rh, cl = oracle_mod.set_density(radii, rhos)
    np.testing.assert_array_equal(prop.radii, rd)
    np.testing.assert_array_equal(prop.rhos, rh)
    np.testing.assert_array_equal(prop.coslimit, cl)
    np.testing.assert_array_equal(prop.maxlayers, oracle_mod.maxlayers(cos, cl))
    if model == "norow0":      # without an r = 0 row the deepest path crosses every row
        assert prop.maxlayers[0] == len(radii)


def test_density_file_loader(tmp_path):
    f = tmp_path / "prem12.dat"
    f.write_text("\n".join(f"{r} {d}" for r, d in W.PREM12) + "\n")
    a, b = Propagator(4, 4), Propagator(4, 4)
    a.setDensityFromFile(str(f))
    b.setDensity(W.PREM12[:, 0], W.PREM12[:, 1])
    np.testing.assert_array_equal(a.radii, b.radii)
    np.testing.assert_array_equal(a.rhos, b.rhos)
    with pytest.raises(RuntimeError, match="could not open density file"):
        a.setDensityFromFile(str(tmp_path / "missing.dat"))


def test_error_conventions():
    p = Propagator(4, 5)
    with pytest.raises(RuntimeError, match="rhos.size\\(\\) != radii.size\\(\\)"):
        p.setDensity([0, 1, 2], [1, 2])
    with pytest.raises(RuntimeError, match="must not be empty"):
        p.setDensity([], [])
    with pytest.raises(RuntimeError, match="radii order messed up"):
        p.setDensity([0, 3000, 2000, 6371], [1, 2, 3, 4])
    with pytest.raises(RuntimeError, match="must set cosine list before production height"):
        p.setProductionHeight(22.0)
    with pytest.raises(RuntimeError, match="number of energy nodes"):
        p.setEnergyList(np.ones(4))
    with pytest.raises(RuntimeError, match="number of cosine nodes"):
        p.setCosineList(np.ones(5))


def test_height_reapplied_on_new_cosines():
    p = Propagator(3, 1)
    p.setCosineList([-1, 0, 1])
    p.setProductionHeight(15.0)
    assert p.ProductionHeightinCentimeter == 15.0 * 100000.0
    p.setCosineList([-0.5, 0, 0.5])
    assert p.ProductionHeightinCentimeter == 15.0 * 100000.0


def test_float_mirror_is_float32():
    p = Propagator(8, 2, dtype=np.float32)
    p.setMNSMatrix(*W.NUFIT[:4])
    p.setNeutrinoMasses(*W.NUFIT[4:])
    p.setCosineList(np.linspace(-1, 1, 8))
    p.setDensity(W.PREM4[:, 0], W.PREM4[:, 1])
    assert p.Mix_U.dtype == np.float32 and p.dm.dtype == np.float32 and p.coslimit.dtype == np.float32
    d64 = Propagator(8, 2)
    d64.setMNSMatrix(*W.NUFIT[:4])
    np.testing.assert_allclose(p.Mix_U, d64.Mix_U, atol=1e-7)
