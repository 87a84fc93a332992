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
    q = np.array(W.NUFIT) * np.array([1.01, 0.99, 1.02, 1.0, 1.01, 0.99])
    ll = chain.loglike(q)
    P = chain.prop.getProbabilities().reshape(-1).astype(np.float64)
    D = chain.data.cpu().numpy().astype(np.float64)
    host = -0.5 * ((P - D) ** 2).sum() / 0.1 ** 2
    assert ll < 0 and abs(ll - host) <= 1e-9 * abs(host)
    st = chain.run(20)
    assert st.steps == 20 and len(st.step_ms) == 20 and 0 <= st.accepted <= 20
    chain.close()
