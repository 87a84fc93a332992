"""Work accounting of the propagation path, used for roofline fractions.

Two counts, both add/sub/mul = 1 and FMA = 2 flops, conversions, selects, rounding (rint),
min/max and bit operations not counted:

* The algorithm this build's mixed-precision pair kernels run (``mixed_flops_per_pair``,
  pair_mixed.hpp): per layer two phase factors (FP64: t = r*L, t - q, (t - q)*pi, the parity
  add; FP32: 22 flops of polynomials and sign for e^{i phi}, 23 for e^{i phi} - 1), the nine
  probabilities from the 2x2 block (22), and for L crossed shells:
    - L = 0: single-grid kernels: the vacuum block from the constants U_xk conj(U_yk) (64);
      batched kernel: the atmosphere columns G (76) and the block from G (28);
    - L = 1: G (76), the innermost shell on G as two rank-1 terms (112) and the block (28);
    - L >= 2, eigenbasis form (two columns carried; single-grid kernels with the basis changes
      from the table, form "table", batched kernel forming them, form "local"): the first shell
      D_1 M_1 (8), per further shell of the way in one basis change and phase scaling of two
      columns (84) and two FP64 phase sums, the innermost basis change (60), the completion of
      Z' (normalised columns 34, cross product 42, times the determinant 18, and the
      determinant's phase factor 22 + 4 FP64), G (76), h = D_L Z' a_x (66), u = Z' G (132) and
      the block (88); form "local" adds per cosine row, or per pair of rows with the same L
      propagated together, 9 flops for M_1 = O_1^T and 54 for each further basis change
      (O_s = [v_1 x v_2, v_1, v_2] and nine dot products);
  FP32 per pair: L = 0: 132 (form "table") or 172 (form "local"), L = 1: 330, L >= 2: 128 L + 446
  (plus the basis changes of form "local"); FP64: 8 (L + 1), plus 2 L + 2 for L >= 2.  This is the roofline's numerator; it agrees within a few per cent with the
  FP32/FP64 flops that rocprofv3's SQ_INSTS_VALU_{ADD,MUL,FMA}_F{32,64} counters report.
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

# pieces of the mixed kernels' count (pair_mixed.hpp), per pair
MIXED_PHASE_F32 = 23          # one phase factor e^{i phi} - 1: polynomials, sign
MIXED_DIAG_F32 = 22           # one phase factor e^{i phi} (eigenbasis form)
MIXED_PHASE_F64 = 4           # t = r L, t - q, (t - q) pi, q + 1.5 2^52
MIXED_ATMOSPHERE = 76         # atmosphere columns G
MIXED_PROBS = 22              # four |T|^2 and five unitarity closures
MIXED_BLOCK_SHALLOW = 28      # the 2x2 block from G for L < 2
MIXED_INNERMOST = 112         # a shell on two columns as two rank-1 terms (apply_layer<2>)
MIXED_VACUUM_BLOCK = 64       # L = 0, single-grid: delta + sum_k w_k U_xk conj(U_yk) for four entries
# eigenbasis form (L >= 2), two columns of Z carried
EIGEN_FIRST_SHELL = 8         # Z = D_1 M_1
EIGEN_SHELL = 84              # Z <- D_i M_i Z
EIGEN_SHELL_F64 = 2           # the determinant's phase sum
EIGEN_INNERMOST = 60          # Z' = M_L Z
EIGEN_COMPLETE = 94           # normalised columns, c_0 x c_1 times the determinant
LOCAL_BASIS_FIRST = 9         # form "local": O_1 and M_1 = O_1^T
LOCAL_BASIS = 54              # form "local": O_s and M_s = O_s^T O_{s-1}
EIGEN_TAIL = 66 + 132 + 88    # h = D_L Z' a_x, u = Z' G, the 2x2 block
LIKELIHOOD_F64_PER_PAIR = 27  # nine (P - D)^2 accumulated in FP64 (likelihood variant)


FORMS = ("table", "local")


def mixed_flops_per_pair(L, form: str = "table"):
    """(FP32, FP64) flops of the mixed pair kernels for a pair crossing L shells, without form
    "local"'s basis changes (counted per row group, local_basis_flops): form "table" for the
    single-grid kernels, "local" for the batched kernel."""
    if form not in FORMS:
        raise ValueError(f"form must be one of {FORMS}")
    L = np.asarray(L, dtype=np.int64)
    shallow = (2 * MIXED_PHASE_F32 + MIXED_ATMOSPHERE + MIXED_BLOCK_SHALLOW + MIXED_PROBS +
               np.where(L == 1, 2 * MIXED_PHASE_F32 + MIXED_INNERMOST, 0))
    if form == "table":
        shallow = np.where(L == 0, 2 * MIXED_PHASE_F32 + MIXED_VACUUM_BLOCK + MIXED_PROBS, shallow)
    deep = (2 * MIXED_PHASE_F32 + 2 * MIXED_DIAG_F32 * L + EIGEN_FIRST_SHELL + EIGEN_SHELL * (L - 2) +
            EIGEN_INNERMOST + EIGEN_COMPLETE + MIXED_DIAG_F32 + MIXED_ATMOSPHERE + EIGEN_TAIL + MIXED_PROBS)
    f32 = np.where(L >= 2, deep, shallow)
    f64 = 2 * MIXED_PHASE_F64 * (L + 1) + np.where(L >= 2, EIGEN_SHELL_F64 * (L - 1) + MIXED_PHASE_F64, 0)
    return f32, f64


def local_basis_flops(maxlayers, n_energies: int) -> float:
    """Form "local"'s basis changes over a grid: the batched kernel takes cosine rows in pairs
    (2j, 2j + 1) and forms them once for a pair whose rows cross the same L >= 2 shells, once per
    row otherwise (an unpaired last row included)."""
    L = np.asarray(maxlayers, dtype=np.int64)
    per_group = lambda x: np.where(x >= 2, LOCAL_BASIS_FIRST + LOCAL_BASIS * (x - 1), 0)
    n = len(L) // 2 * 2
    a, b = L[0:n:2], L[1:n:2]
    groups = np.where(a == b, per_group(a), per_group(a) + per_group(b)).sum()
    if len(L) % 2:
        groups += per_group(L[-1])
    return float(groups) * n_energies


def mixed_grid_flops(maxlayers, n_energies: int, likelihood: bool = False,
                     form: str = "table") -> tuple[float, float]:
    """(FP32, FP64) flops of one mixed pair-kernel launch over a grid."""
    f32, f64 = mixed_flops_per_pair(maxlayers, form)
    n_pairs = len(np.asarray(maxlayers)) * n_energies
    extra = LIKELIHOOD_F64_PER_PAIR * n_pairs if likelihood else 0
    local = local_basis_flops(maxlayers, n_energies) if form == "local" else 0.0
    return float(f32.sum()) * n_energies + local, float(f64.sum()) * n_energies + extra


def f64_hoisted_flops_per_pair(L):
    """FP64 flops of the hoisted FP64 pair kernel (propagate_pairs.hip pairs_f64_kernel) for a
    pair crossing L shells, as written (add/sub/mul/div = 1; sincos excluded): per layer the
    three phase arguments r L / E (6) and A = sum_k e^{i phi_k} Q_k from the packed Hermitian
    projectors (138); the way-in product fin = A fin for layers 1..L and the mantle product for
    layers 2..L-1 and the final mantle fin for L >= 2 (198 each, a 3x3 complex product); the nine
    |amplitude|^2 (27).  F(0) = 171, F(1) = 513, F(L >= 2) = 540 L - 27."""
    L = np.asarray(L, dtype=np.int64)
    return np.where(L == 0, 171, np.where(L == 1, 513, 540 * L - 27))


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
