"""Work accounting of the propagation path, used for roofline fractions.

Two counts, both add/sub/mul = 1 and FMA = 2 flops, conversions, selects, rounding (rint),
min/max and bit operations not counted:

* The algorithm this build's mixed-precision pair kernel runs (``mixed_flops_per_pair``,
  pair_mixed.hpp): per layer two phase factors (FP64: t = r*L, t - q, (t - q)*pi, the parity
  add; FP32: 23 flops of polynomials and sign), per shell of the way in one update of the 3x3
  running product by two real rank-1 terms (168), and once per pair the first shell (54), the
  atmosphere (76), the product with the atmosphere columns (132), the innermost shell on two
  columns (112), the 2x2 block (130; 28 without a running product) and the nine probabilities
  (22).  For L crossed shells:
      FP32: L = 0: 172, L = 1: 330, L >= 2: 214 L + 236;   FP64: 8 (L + 1).
  This is the roofline's numerator: it agrees within about 1 % with the FP32/FP64 flops that
  rocprofv3's SQ_INSTS_VALU_{ADD,MUL,FMA}_F{32,64} counters report for the kernel.
* The reference algorithm as written (physics.hpp:268-702), kept as a measure of the work the
  reference would do for the same results (``reference_work``): one layer step
  (get_transition_matrix) = getMfast 69 + get_product 788 + getA 874 = 1731 flops (plus 2 sqrt,
  1 acos, 3 cos, 3 sincos); a pair crossing L shells runs L+1 layer steps, 2L-1 running-product
  multiplies in the loop (0 for L = 0) and one final multiply (216 flops each), and 27 flops for
  the nine |A|^2: F(0) = 1974, F(L >= 1) = 2163 L + 1758.  The mixed kernel computes the same
  probabilities with about a tenth of this (the eigen-decomposition hoisted to one per
  (shell, energy), rank-1 updates in a real basis, unitarity closure).

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

# pieces of the mixed kernel's count (pair_mixed.hpp), per pair
MIXED_PHASE_F32 = 23          # one phase factor e^{i phi} - 1: polynomials, sign
MIXED_PHASE_F64 = 4           # t = r L, t - q, (t - q) pi, q + 1.5 2^52
MIXED_SHELL_UPDATE = 168      # 3x3 running product, two rank-1 terms (apply_layer<3>)
MIXED_FIRST_SHELL = 54        # init_layer
MIXED_ATMOSPHERE = 76         # atmosphere columns G
MIXED_Y_TIMES_G = 132         # running product times G
MIXED_INNERMOST = 112         # innermost shell on two columns (apply_layer<2>)
MIXED_BLOCK = 130             # the 2x2 block through W Y^T
MIXED_BLOCK_SHALLOW = 28      # the 2x2 block for L < 2
MIXED_PROBS = 22              # four |T|^2 and five unitarity closures
LIKELIHOOD_F64_PER_PAIR = 27  # nine (P - D)^2 accumulated in FP64 (likelihood variant)


def mixed_flops_per_pair(L):
    """(FP32, FP64) flops of the mixed pair kernel for a pair crossing L shells."""
    L = np.asarray(L, dtype=np.int64)
    phases = 2 * (L + 1)
    f32 = phases * MIXED_PHASE_F32 + MIXED_ATMOSPHERE + MIXED_PROBS
    f32 = f32 + np.where(L >= 1, MIXED_INNERMOST, 0)
    f32 = f32 + np.where(L >= 2, MIXED_FIRST_SHELL + MIXED_SHELL_UPDATE * (L - 2) + MIXED_Y_TIMES_G + MIXED_BLOCK,
                         MIXED_BLOCK_SHALLOW)
    f64 = phases * MIXED_PHASE_F64
    return f32, f64


def mixed_grid_flops(maxlayers, n_energies: int, likelihood: bool = False) -> tuple[float, float]:
    """(FP32, FP64) flops of one mixed pair-kernel launch over a grid."""
    f32, f64 = mixed_flops_per_pair(maxlayers)
    n_pairs = len(np.asarray(maxlayers)) * n_energies
    extra = LIKELIHOOD_F64_PER_PAIR * n_pairs if likelihood else 0
    return float(f32.sum()) * n_energies, float(f64.sum()) * n_energies + extra


def valu_roofline(f32_flops: float, f64_flops: float, kernel_ms: float) -> dict:
    """Roofline entries of a vector-ALU kernel: achieved = all flops / time against the FP32
    vector peak (FP64 flops are counted at the FP32 rate, which understates the fraction by
    their share), with the mixed formula (F32 / 157.3 + F64 / 78.6) / t and the fraction of the
    non-packed FP32 issue rate beside it."""
    t = kernel_ms * 1e-3
    achieved = (f32_flops + f64_flops) / t / 1e12
    return {"achieved": achieved, "peak": PEAK_FP32_VALU_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_VALU_TFLOPS,
            "frac_mixed_formula": (f32_flops / (PEAK_FP32_VALU_TFLOPS * 1e12) +
                                   f64_flops / (PEAK_FP64_VALU_TFLOPS * 1e12)) / t,
            "frac_of_nonpacked_fp32_peak": achieved / PEAK_FP32_NONPACKED_TFLOPS,
            "flops_per_launch": {"f32": f32_flops, "f64": f64_flops}}


def flops_per_pair(L):
    """Reference-algorithm flops of a pair crossing L shells."""
    L = np.asarray(L, dtype=np.int64)
    return np.where(L == 0, 1974, 2163 * L + 1758)


def grid_flops(maxlayers, n_energies: int) -> float:
    """Reference-algorithm flops of one propagation of a grid (sum over cosines)."""
    return float(flops_per_pair(maxlayers).sum()) * n_energies


def maxlayers_of(cosines, radii, rhos):
    """maxlayers exactly as the host set-up computes them (propagator.hpp:117-161, 301-307)."""
    from .propagator import Propagator
    p = Propagator(len(cosines), 1)
    p.setCosineList(cosines)
    p.setDensity(radii, rhos)
    return p.maxlayers.copy()


# Surrogate network, flops per evaluated row of the 9 -> 256 x 5 -> 9 architecture, counted
# once per product (the algorithmic work): 2 * (9*256 + 4*256*256 + 256*9).
MLP_FLOPS_PER_ROW = 2 * (9 * 256 + 4 * 256 * 256 + 256 * 9)
# what the kernels execute: every input and activation enters the products as two bf16 terms
# (hi + lo): 2 * (18*256 + 4*512*256 + 512*9)
MLP_EXECUTED_FLOPS_PER_ROW = 2 * (18 * 256 + 4 * 512 * 256 + 512 * 9)
