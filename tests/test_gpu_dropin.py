"""The C++ drop-in API (include/cudaprob3/hippropagator.hpp) on the GPU: the reference
example's call sequence through std::unique_ptr<Propagator<FLOAT_T>>, results read back
with getProbability, compared with the oracle."""
import os
import subprocess

import numpy as np
import pytest

from cudaprob3_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_bin", "dropin")


def run_dropin(tmp_path, cos, E, model, params, h_km, antinu, type_, precision):
    dens = tmp_path / "model.dat"
    dens.write_text("\n".join(f"{float(r)!r} {float(d)!r}" for r, d in model) + "\n")
    out = tmp_path / "P.bin"
    spec = tmp_path / "spec.txt"
    lines = [f"{len(cos)} {len(E)} {antinu} {float(h_km)!r} {type_} {precision}",
             " ".join(repr(float(x)) for x in params),
             " ".join(repr(float(x)) for x in cos),
             " ".join(repr(float(x)) for x in E),
             str(dens), str(out)]
    spec.write_text("\n".join(lines) + "\n")
    print(f"[gpu] dropin {type_} precision={precision} {len(cos)}x{len(E)}", flush=True)
    r = subprocess.run([BIN, str(spec)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return np.fromfile(out, dtype=np.float64).reshape(len(cos), len(E), 9)


@pytest.mark.parametrize("type_,precision,tol", [("f64", 0, 1e-12), ("f64", 1, 1e-12), ("f64", 2, 1e-5),
                                                 ("f32", 2, 5e-4)])
def test_cpp_dropin_matches_oracle(tmp_path, oracle_mod, type_, precision, tol):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/_bin/dropin is not built (run __graft_entry__.build())")
    cos = np.linspace(-1, 1, 40)
    E = np.geomspace(0.1, 100, 70)
    model = W.PREM12
    P = run_dropin(tmp_path, cos, E, model, W.NUFIT, 22.0, 1, type_, precision)
    if type_ == "f32":
        cos, E = cos.astype(np.float32), E.astype(np.float32)
        model = model.astype(np.float32)
        params = [float(np.float32(x)) for x in W.NUFIT]
    else:
        params = W.NUFIT
    ref = oracle_mod.propagate(1, cos, E, model[:, 0], model[:, 1], 22.0, params)
    assert np.abs(P - ref).max() <= tol
