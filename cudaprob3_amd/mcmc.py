"""Markov-chain parameter fit over a fixed (cosZ, E) grid on one MI355X.

Each step proposes new oscillation parameters (Gaussian, sigma = 1 % of each parameter's
range), re-propagates the whole grid with the mixed-precision kernel fused with the comparison
against pseudo-data (sum of squared differences in FP64 per workgroup, cp3h_compute_sqdiff_f32:
the probabilities are never written to HBM) and accepts or rejects with the Metropolis rule.
Only one double per step crosses PCIe.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from . import workloads as W
from .propagator import HipPropagatorSingle

PARAM_RANGE = W.PARAM_HI - W.PARAM_LO


@dataclass
class ChainStats:
    steps: int = 0
    accepted: int = 0
    step_ms: list = field(default_factory=list)
    loglike: list = field(default_factory=list)

    def percentile(self, q):
        return float(np.percentile(self.step_ms, q)) if self.step_ms else float("nan")


class Chain:
    def __init__(self, grid: W.Grid, true_params, sigma_p: float = 0.01, seed: int = 5, device: int = 0):
        import torch   # device memory for the pseudo-data only

        self.grid = grid
        self.sigma_p = sigma_p
        self.rng = np.random.Generator(np.random.PCG64(seed))
        self.prop = HipPropagatorSingle(device, len(grid.cosines), len(grid.energies), precision="mixed")
        p = self.prop
        p.setEnergyList(grid.energies)
        p.setCosineList(grid.cosines)
        p.setDensity(grid.radii, grid.rhos)
        p.setProductionHeight(grid.h_km)
        self._lib = _lib.load()
        n = 9 * len(grid.cosines) * len(grid.energies)
        self.data = torch.empty(n, dtype=torch.float32, device=f"cuda:{device}")
        torch.cuda.current_stream(self.data.device).synchronize()   # written by the propagator's stream
        self._evaluate(true_params)
        _lib.check(self._lib.cp3h_copy_results_device(p._handle(), ctypes.c_void_p(self.data.data_ptr()), n))
        self.params = np.array(true_params, dtype=np.float64)
        self.current = self.loglike(self.params)
        self.stats = ChainStats()

    @property
    def n_pairs(self):
        return len(self.grid.cosines) * len(self.grid.energies)

    def _evaluate(self, params):
        p = self.prop
        p.setMNSMatrix(*params[:4])
        p.setNeutrinoMasses(*params[4:])
        p.calculateProbabilitiesAsync(self.grid.antinu)

    def loglike(self, params) -> float:
        """-1/2 sum (P - D)^2 / sigma_P^2 over all nine channels and every pair (fused kernel)."""
        p = self.prop
        p.setMNSMatrix(*params[:4])
        p.setNeutrinoMasses(*params[4:])
        return -0.5 * p.computeSqdiff(self.grid.antinu, self.data.data_ptr()) / self.sigma_p ** 2

    def loglike_stored(self, params) -> float:
        """The same through stored results (calculateProbabilities, then the device comparison)."""
        self._evaluate(params)
        r = ctypes.c_double()
        _lib.check(self._lib.cp3h_sqdiff_results_f32(self.prop._handle(), ctypes.c_void_p(self.data.data_ptr()),
                                                     ctypes.byref(r)))
        return -0.5 * r.value / self.sigma_p ** 2

    def propose(self):
        q = self.params + self.rng.normal(0.0, 0.01 * PARAM_RANGE)
        q[3] = math.fmod(q[3], 2 * math.pi)
        if q[3] < 0:
            q[3] += 2 * math.pi
        return q

    def step(self):
        t0 = time.perf_counter()
        q = self.propose()
        ll = self.loglike(q)
        accept = math.log(self.rng.random() + 1e-300) < ll - self.current
        if accept:
            self.params, self.current = q, ll
            self.stats.accepted += 1
        self.stats.steps += 1
        self.stats.step_ms.append((time.perf_counter() - t0) * 1e3)
        self.stats.loglike.append(self.current)
        return accept

    def run(self, steps: int):
        for _ in range(steps):
            self.step()
        return self.stats

    def close(self):
        self.prop.close()


class BinnedChain:
    """Metropolis chain on bin-averaged probabilities: every step evaluates the whole grid of
    cosZ x energy bins (the grid's energies taken as bin centres, +-2.5 % bins, 16-point
    Gauss-Legendre) either with the exact mixed kernel alone (``method="exact"``, binned in
    registers) or with the surrogate network plus its exact fallback (``method="surrogate"``),
    then compares with pseudo-data made by the same method at the true parameters (FP64 sum on
    the device).  The surrogate must have been trained for the grid's Earth model."""

    def __init__(self, grid: W.Grid, true_params, method: str, net=None, emap=None, earth: str = "fine100",
                 sigma_p: float = 0.01, seed: int = 5, device: int = 0):
        import torch
        from . import surrogate as S
        self.torch = torch
        self.grid, self.sigma_p, self.method = grid, sigma_p, method
        self.rng = np.random.Generator(np.random.PCG64(seed))
        self.antinu = int(grid.antinu)
        if method == "exact":
            self.ev = S.ExactBinnedGrid(torch, grid.cosines, grid.energies, slots=1, device=device, earth=earth)
        elif method == "surrogate":
            k = S.hypothesis_class(true_params, self.antinu)
            self.ev = S.GridEvaluator(net, emap, grid.cosines, grid.energies, slots=1, device=device, earth=earth,
                                      classes=[k])
            self.exact_fraction = self.ev.fallback_fraction(k)
        else:
            raise ValueError("method must be 'exact' or 'surrogate'")
        self.data = None
        self.data = self._binned(true_params).clone()
        self.params = np.array(true_params, dtype=np.float64)
        self.current = self.loglike(self.params)
        self.stats = ChainStats()

    @property
    def n_bins(self):
        return len(self.grid.cosines) * len(self.grid.energies)

    def _binned(self, params):
        """Device tensor [9][n_cos][n_bins] of the method's binned probabilities (valid until the
        next evaluation)."""
        got = {}
        self.ev.run([tuple(params)], [self.antinu], consume=lambda idx, out: got.update(P=out[0]))
        return got["P"]

    def loglike(self, params) -> float:
        P = self._binned(params)
        if self.data is None:
            return 0.0
        d = P - self.data
        return -0.5 * float(self.torch.sum(d * d, dtype=self.torch.float64)) / self.sigma_p ** 2

    def step(self):
        t0 = time.perf_counter()
        q = self.params + self.rng.normal(0.0, 0.01 * PARAM_RANGE)
        q[3] = math.fmod(q[3], 2 * math.pi) % (2 * math.pi)
        ll = self.loglike(q)
        accept = math.log(self.rng.random() + 1e-300) < ll - self.current
        if accept:
            self.params, self.current = q, ll
            self.stats.accepted += 1
        self.stats.steps += 1
        self.stats.step_ms.append((time.perf_counter() - t0) * 1e3)
        self.stats.loglike.append(self.current)
        return accept

    def run(self, steps: int):
        for _ in range(steps):
            self.step()
        return self.stats

    def close(self):
        self.ev.close()
