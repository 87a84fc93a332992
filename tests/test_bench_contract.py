"""bench.py's contract pieces that run without a GPU: the single-GPU refusal, the command line,
and the algorithmic work accounting its roofline divides by."""
import os
import subprocess
import sys

import numpy as np

from cudaprob3_amd import metrics as M
from cudaprob3_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "single MI355X" in r.stderr and r.stdout == ""


def test_help_lists_the_workloads():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0
    for w in ("jitter12", "scan1024", "mcmc8192", "surrogate2048"):
        assert w in r.stdout


def test_flops_per_pair_from_its_parts():
    """F(L) from the reference's structure (physics.hpp:645-699): L+1 layer steps of 1731 flops,
    2L-1 running-product multiplies in the loop plus the final one (L >= 1) or none plus the
    final one (L = 0), 216 flops each, and 27 flops for the nine |A|^2."""
    L = np.arange(0, 120)
    mults = np.where(L == 0, 1, 2 * L)
    expect = M.FLOPS_LAYER_STEP * (L + 1) + M.FLOPS_MATMUL * mults + 27
    assert np.array_equal(M.flops_per_pair(L), expect)


def test_jitter12_grid_flops():
    """The headline grid's algorithmic work per launch (the roofline's numerator): 34.60 GFLOP,
    from the host-computed maxlayers of the jittered 12-shell Earth."""
    g = W.jitter12_2048()
    ml = M.maxlayers_of(g.cosines, g.radii, g.rhos)
    assert ml.max() == 12 and (ml == 0).sum() == 1024
    assert abs(M.grid_flops(ml, len(g.energies)) / 1e9 - 34.60) < 0.005


def _profile(name):
    import json
    with open(os.path.join(ROOT, "profiles", name)) as f:
        return json.load(f)


def test_mixed_flops_agree_with_the_committed_counters():
    """The roofline's numerator for the single-grid mixed kernels (eigenbasis form, vacuum rows,
    the likelihood's FP64 sums) against the flops rocprofv3's SQ_INSTS_VALU_{ADD,MUL,FMA}_F{32,64}
    counters read per launch (profiles/traffic_*_mixed.json): within 3 %.  On the 12-shell grid a
    count without the completion of Z' (normalisation, cross product, determinant; about 10 % of
    a deep pair's flops there) would miss by more than that."""
    g = W.jitter12_2048()
    ml = M.maxlayers_of(g.cosines, g.radii, g.rhos)
    f32, f64 = M.mixed_grid_flops(ml, len(g.energies))
    counted = _profile("traffic_jitter12_2048_mixed.json")["counter_flops_per_launch"]
    assert abs((f32 + f64) / counted - 1) < 0.03
    g = W.fine100_grid(8192)
    ml = M.maxlayers_of(g.cosines, g.radii, g.rhos)
    f32, f64 = M.mixed_grid_flops(ml, len(g.energies), likelihood=True)
    counted = _profile("traffic_mcmc8192_mixed.json")["counter_flops_per_launch"]
    assert abs((f32 + f64) / counted - 1) < 0.03
    # the batched kernel's form (no vacuum rows, basis changes formed per row group) does not
    # describe this kernel
    f32b, f64b = M.mixed_grid_flops(ml, len(g.energies), likelihood=True, form="local")
    assert abs((f32b + f64b) / counted - 1) > 0.03


def test_batched_flops_agree_with_the_committed_counters():
    """The batched kernel (eigenbasis form, basis changes formed per row group) over the 1024 scan
    sets: the counters also count the lanes
    of each set's last 64-energy tile beyond its energies (masked lanes still issue), so the
    algorithmic count with energies padded to 64 must match them within 3 %; the roofline uses the
    unpadded count (useful work only), about 5 % lower."""
    useful = padded = 0.0
    for s in W.scan_sets():
        ml = M.maxlayers_of(s.cosines, s.radii, s.rhos)
        a, b = M.mixed_grid_flops(ml, len(s.energies), form="local")
        useful += a + b
        a, b = M.mixed_grid_flops(ml, -(-len(s.energies) // 64) * 64, form="local")
        padded += a + b
    counted = _profile("traffic_scan1024_mixed.json")["counter_flops_per_launch"]
    assert abs(padded / counted - 1) < 0.03
    assert useful < padded
