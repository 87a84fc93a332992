"""The allocation-free entry point cp3h_propagate_mixed_noalloc against the handle-based mixed
propagation (cp3h_create / cp3h_compute on a CP3H_MIXED handle), which runs the same set-up
and pair kernels: results must be bitwise equal, on two calls in a row (the second one smaller,
on another Earth model and hypothesis, so it reuses scratch the first one filled)."""
import ctypes

import numpy as np
import pytest

from cudaprob3_amd import _lib
from cudaprob3_amd import workloads as W
from cudaprob3_amd.propagator import HipPropagatorSingle, Propagator


def handle_result(cos, E, radii, rhos, params, h_km, antinu):
    with HipPropagatorSingle(0, len(cos), len(E), precision="mixed") as p:
        p.setEnergyList(E)
        p.setCosineList(cos)
        p.setMNSMatrix(*params[:4])
        p.setNeutrinoMasses(*params[4:])
        p.setDensity(radii, rhos)
        p.setProductionHeight(h_km)
        p.calculateProbabilities(antinu)
        return p.getProbabilities().astype(np.float32)          # float results, exact in float64


def noalloc_result(torch, cos, E, radii, rhos, params, h_km, antinu):
    host = Propagator(len(cos), len(E))                          # the reference's host set-up
    host.setEnergyList(E)
    host.setCosineList(cos)
    host.setMNSMatrix(*params[:4])
    host.setNeutrinoMasses(*params[4:])
    host.setDensity(radii, rhos)
    host.setProductionHeight(h_km)
    dev = torch.device("cuda:0")
    d = lambda a, t=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=t, device=dev)
    cos_d, ml_d, e_d = d(host.cosineList), d(host.maxlayers, torch.int32), d(host.energyList)
    r_d, rho_d = d(host.radii), d(host.rhos)
    out = torch.full((9, len(cos), len(E)), -1.0, dtype=torch.float32, device=dev)
    U = np.ascontiguousarray(host.Mix_U, dtype=np.float64)
    D = np.ascontiguousarray(host.dm, dtype=np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    stream = torch.cuda.current_stream()
    print(f"[gpu] noalloc {len(cos)}x{len(E)} rows={len(radii)} antinu={antinu}", flush=True)
    lib = _lib.load()
    _lib.check(lib.cp3h_propagate_mixed_noalloc(
        0, ctypes.c_void_p(cos_d.data_ptr()), ctypes.c_void_p(ml_d.data_ptr()), len(cos),
        ctypes.c_void_p(e_d.data_ptr()), len(E), ctypes.c_void_p(r_d.data_ptr()), ctypes.c_void_p(rho_d.data_ptr()),
        len(radii), U.ctypes.data_as(dp), D.ctypes.data_as(dp), int(antinu),
        float(host.ProductionHeightinCentimeter), ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream.cuda_stream)))
    stream.synchronize()
    return out.cpu().numpy()


@pytest.mark.gpu
def test_noalloc_matches_handle_path_twice():
    torch = pytest.importorskip("torch")
    assert torch.cuda.is_available()
    g = W.jitter12_2048()
    cases = [
        # first call: 700 cosines x 900 energies of the 12-shell Earth, nubar
        (g.cosines[::3][:700], g.energies[::2][:900], g.radii, g.rhos, g.params, g.h_km, 1),
        # second call, smaller: 37 x 65 on the 4-shell Earth, another hypothesis, nu
        (np.linspace(-1, 1, 37), np.geomspace(0.1, 100, 65), W.PREM4[:, 0], W.PREM4[:, 1],
         [0.57, 0.15, 0.8, 1.2, 7.0e-5, -2.45e-3], 15.0, 0),
    ]
    for case in cases:
        ref = handle_result(*case)
        got = noalloc_result(torch, *case)
        assert got.shape == ref.shape
        np.testing.assert_array_equal(got, ref)


def test_noalloc_refuses_oversized_grid():
    """Host-side refusal before any device call (runs without a GPU)."""
    lib = _lib.load()
    dp = ctypes.POINTER(ctypes.c_double)
    U = np.zeros(18)
    D = np.zeros(9)
    one = ctypes.c_void_p(16)          # never dereferenced: the size check comes first
    rc = lib.cp3h_propagate_mixed_noalloc(0, one, one, 8192, one, 8192, one, one, 200, U.ctypes.data_as(dp),
                                          D.ctypes.data_as(dp), 0, 0.0, one, None)
    assert rc == 1 and "too large" in lib.cp3h_last_error().decode()
