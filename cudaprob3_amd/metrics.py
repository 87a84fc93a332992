"""Work accounting of the propagation path, used for roofline fractions.

Algorithmic floating-point operations follow the reference algorithm as written
(physics.hpp:268-702), counting add/sub/mul/div as 1 and an FMA as 2, transcendental
calls excluded:

* one layer step (get_transition_matrix) = getMfast 69 + get_product 788 + getA 874
  = 1731 flops, plus 2 sqrt, 1 acos, 3 cos and 3 sincos;
* a pair crossing L shells runs L+1 layer steps, 2L-1 running-product multiplies in the
  loop (0 for L = 0) and one final multiply (216 flops each), and 27 flops for the nine
  |A|^2: F(0) = 1974, F(L >= 1) = 2163 L + 1758.

Output bytes are 9 values per pair: 36 B in FP32, 72 B in FP64.
"""
from __future__ import annotations

import numpy as np

FLOPS_LAYER_STEP = 1731
FLOPS_MATMUL = 216

# MI355X published peaks
PEAK_FP32_VALU_TFLOPS = 157.3      # vector FP32 (with packed FP32)
PEAK_FP32_NONPACKED_TFLOPS = 78.6  # vector FP32 without packed instructions (disabled in this build)
PEAK_FP64_VALU_TFLOPS = 78.6       # vector FP64
PEAK_HBM_GBS = 8000.0
PEAK_BF16_MFMA_TFLOPS = 2500.0


def flops_per_pair(L):
    L = np.asarray(L, dtype=np.int64)
    return np.where(L == 0, 1974, 2163 * L + 1758)


def grid_flops(maxlayers, n_energies: int) -> float:
    """Algorithmic flops of one propagation of a grid (sum over cosines)."""
    return float(flops_per_pair(maxlayers).sum()) * n_energies


def maxlayers_of(cosines, radii, rhos):
    """maxlayers exactly as the host set-up computes them (propagator.hpp:117-161, 301-307)."""
    from .propagator import Propagator
    p = Propagator(len(cosines), 1)
    p.setCosineList(cosines)
    p.setDensity(radii, rhos)
    return p.maxlayers.copy()


# Surrogate network, algorithmic flops per evaluated row: the forward products with two-term
# (hi + lo) inputs and activations, without padding: 2 * (18*256 + 4*512*256 + 512*9).
MLP_FLOPS_PER_ROW = 2 * (18 * 256 + 4 * 512 * 256 + 512 * 9)
