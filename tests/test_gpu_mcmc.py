"""Device-side likelihood and the Markov-chain driver on a small 100-shell grid."""
import ctypes

import numpy as np
import pytest

from cudaprob3_amd import workloads as W
from cudaprob3_amd.mcmc import Chain

pytestmark = pytest.mark.gpu


def test_device_sqdiff_matches_host():
    import torch
    grid = W.fine100_grid(48)
    chain = Chain(grid, W.NUFIT, sigma_p=0.1, seed=5)
    print("[gpu] mcmc small grid", flush=True)
    assert chain.loglike(np.array(W.NUFIT)) == 0.0         # data were made at these parameters
    assert chain.loglike_stored(np.array(W.NUFIT)) == 0.0
    q = np.array(W.NUFIT) * np.array([1.01, 0.99, 1.02, 1.0, 1.01, 0.99])
    ll_fused = chain.loglike(q)
    ll = chain.loglike_stored(q)
    P = chain.prop.getProbabilities().reshape(-1).astype(np.float64)
    D = chain.data.cpu().numpy().astype(np.float64)
    host = -0.5 * ((P - D) ** 2).sum() / 0.1 ** 2
    assert ll < 0 and abs(ll - host) <= 1e-9 * abs(host)
    assert abs(ll_fused - ll) <= 1e-12 * abs(ll)         # same P, same FP64 terms, other order
    st = chain.run(20)
    assert st.steps == 20 and len(st.step_ms) == 20 and 0 <= st.accepted <= 20
    chain.close()


def test_fused_likelihood_ragged_grid():
    """Grids that do not fill whole wavefronts or workgroups (idle lanes add nothing)."""
    import torch
    from cudaprob3_amd.propagator import HipPropagatorSingle
    print("[gpu] fused likelihood, ragged grid", flush=True)
    grid = W.fine100_grid(48)
    cos, E = grid.cosines[:13], grid.energies[:37]
    p = HipPropagatorSingle(0, len(cos), len(E), precision="mixed")
    p.setEnergyList(E)
    p.setCosineList(cos)
    p.setDensity(grid.radii, grid.rhos)
    p.setProductionHeight(grid.h_km)
    p.setMNSMatrix(*W.NUFIT[:4])
    p.setNeutrinoMasses(*W.NUFIT[4:])
    rng = np.random.default_rng(3)
    D = rng.uniform(0, 1, 9 * len(cos) * len(E)).astype(np.float32)
    d = torch.as_tensor(D, device="cuda")
    torch.cuda.synchronize()
    s = p.computeSqdiff(1, d.data_ptr())
    p.calculateProbabilities(1)
    P = p.getProbabilities().reshape(-1).astype(np.float64)
    assert abs(s - ((P - D.astype(np.float64)) ** 2).sum()) <= 1e-12 * s
    p.close()


def test_binned_chains_small_grid():
    """Binned chains (exact kernel alone, surrogate + fallback) on a 64 x 64 grid of the 100-shell
    Earth: zero log-likelihood at the true parameters, finite steps, and the two methods' binned
    grids within 1e-3 of each other at the truth."""
    import torch
    from cudaprob3_amd import train_surrogate as T
    from cudaprob3_amd.mcmc import BinnedChain
    print("[gpu] binned chains, 64 x 64", flush=True)
    grid = W.fine100_grid(64)
    net, emap, _ = T.load(T.path_for("fine100"), max_batch=1024)
    truth = {}
    for method in ("exact", "surrogate"):
        ch = BinnedChain(grid, W.NUFIT, method, net=net, emap=emap, earth="fine100", sigma_p=0.1, seed=3)
        assert ch.loglike(np.array(W.NUFIT)) == 0.0
        truth[method] = ch.data.clone()
        st = ch.run(3)
        assert st.steps == 3 and all(np.isfinite(st.loglike))
        ch.close()
    assert float((truth["surrogate"] - truth["exact"]).abs().max()) <= 1e-3
    net.close()
