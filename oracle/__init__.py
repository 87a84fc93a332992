"""CPU oracle for the propagation path -- TEST INFRASTRUCTURE ONLY.

ctypes bindings for ``liboracle.so`` (the project's C++ restatement of the reference's
``CpuPropagator<double>`` path, ``cp3_oracle.cpp``) and, when it has been built, for
``_ref/libcp3ref.so`` (the reference itself, compiled from its headers in place).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py``
import this package.  The shipped library never calls it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libcp3ref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_dp)


def build(ref_dir: str | None = "/root/reference") -> None:
    """Compile the oracle (and the reference shim when the reference tree exists)."""
    subprocess.run(["make", "-s", "-B", "-C", HERE, "all"], check=True)
    if ref_dir and os.path.isdir(ref_dir):
        subprocess.run(["make", "-s", "-B", "-C", HERE, "ref", f"REF_DIR={ref_dir}"], check=True)


_lib_cache: dict[str, ctypes.CDLL] = {}


def _load(path: str) -> ctypes.CDLL:
    if path not in _lib_cache:
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run oracle.build())")
        lib = ctypes.CDLL(path)
        _lib_cache[path] = lib
    return _lib_cache[path]


def oracle_lib() -> ctypes.CDLL:
    return _load(ORACLE_SO)


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib() -> ctypes.CDLL:
    return _load(REF_SO)


def _propagate(fn, antinu, cosines, energies, radii, rhos, h_km, params, threads):
    cosines = np.ascontiguousarray(cosines, dtype=np.float64)
    energies = np.ascontiguousarray(energies, dtype=np.float64)
    radii = np.ascontiguousarray(radii, dtype=np.float64)
    rhos = np.ascontiguousarray(rhos, dtype=np.float64)
    n_c, n_e = len(cosines), len(energies)
    out = np.empty((n_c, n_e, 9), dtype=np.float64)
    t12, t13, t23, dcp, dm12, dm23 = (float(x) for x in params)
    fn.restype = ctypes.c_int
    st = fn(ctypes.c_int(int(antinu)), _d(cosines), ctypes.c_int(n_c), _d(energies), ctypes.c_int(n_e),
            _d(radii), _d(rhos), ctypes.c_int(len(radii)), ctypes.c_double(h_km),
            ctypes.c_double(t12), ctypes.c_double(t13), ctypes.c_double(t23), ctypes.c_double(dcp),
            ctypes.c_double(dm12), ctypes.c_double(dm23), ctypes.c_int(threads),
            out.ctypes.data_as(_dp))
    if st != 0:
        raise RuntimeError(f"propagation set-up failed with status {st}")
    return out


def propagate(antinu, cosines, energies, radii, rhos, h_km, params, threads=0):
    """Oracle end to end.  Returns P[ic, ie, t] (AoS, t = in*3+out) in float64."""
    return _propagate(oracle_lib().orc_propagate, antinu, cosines, energies, radii, rhos, h_km, params, threads)


JITTER_SO = os.path.join(HERE, "liboracle_jitter.so")


def libm_sensitivity(antinu, cosines, energies, radii, rhos, h_km, params, threads=0, reference=None):
    """Largest |dP| by which the reference algorithm's FP64 result moves when its per-pair
    sin/cos/acos results are each moved by up to one ulp (another faithfully rounded libm).
    This is the reference's own rounding noise floor for a parity comparison."""
    if reference is None:
        reference = propagate(antinu, cosines, energies, radii, rhos, h_km, params, threads)
    moved = _propagate(_load(JITTER_SO).orc_propagate, antinu, cosines, energies, radii, rhos, h_km, params, threads)
    return float(np.abs(moved - reference).max())


def fp64_tolerance(sensitivity: float, floor: float = 1e-12) -> float:
    """FP64 parity bar: 1e-12, or twice the reference's own libm noise where that is larger."""
    return max(floor, 2.0 * sensitivity)


def ref_propagate(antinu, cosines, energies, radii, rhos, h_km, params, threads=1):
    """The compiled reference CpuPropagator<double>.  Same layout as :func:`propagate`."""
    return _propagate(ref_lib().ref_propagate, antinu, cosines, energies, radii, rhos, h_km, params, threads)


def mns_matrix(t12, t13, t23, dcp):
    U = np.empty(18)
    oracle_lib().orc_mns_matrix(ctypes.c_double(t12), ctypes.c_double(t13), ctypes.c_double(t23),
                                ctypes.c_double(dcp), _d_out(U))
    return U


def mass_matrix(dm12, dm23):
    D = np.empty(9)
    oracle_lib().orc_mass_matrix(ctypes.c_double(dm12), ctypes.c_double(dm23), _d_out(D))
    return D


def _d_out(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def set_density(radii, rhos):
    """Returns (radii_desc, rhos_desc, coslimit) or raises on the reference's error cases."""
    radii = np.ascontiguousarray(radii, dtype=np.float64)
    rhos = np.ascontiguousarray(rhos, dtype=np.float64)
    n = len(radii)
    if len(rhos) != n:
        raise RuntimeError("setDensity : rhos.size() != radii.size()")
    rd, rh, cl = np.empty(max(n, 1)), np.empty(max(n, 1)), np.empty(max(n, 1))
    oracle_lib().orc_set_density.restype = ctypes.c_int
    st = oracle_lib().orc_set_density(_d(radii), _d(rhos), ctypes.c_int(n), _d_out(rd), _d_out(rh), _d_out(cl))
    if st == 2:
        raise RuntimeError("setDensity : vectors must not be empty")
    if st == 3:
        raise RuntimeError("radii order messed up")
    return rd[:n], rh[:n], cl[:n]


def maxlayers(cosines, coslimit):
    cosines = np.ascontiguousarray(cosines, dtype=np.float64)
    coslimit = np.ascontiguousarray(coslimit, dtype=np.float64)
    out = np.empty(len(cosines), dtype=np.int32)
    oracle_lib().orc_maxlayers(_d(cosines), ctypes.c_int(len(cosines)), _d(coslimit), ctypes.c_int(len(coslimit)),
                               out.ctypes.data_as(_ip))
    return out


def vacuum_order(DM9):
    out = np.empty(3, dtype=np.int32)
    oracle_lib().orc_vacuum_order(_d(DM9), out.ctypes.data_as(_ip))
    return out


def matter_masses(U18, DM9, E, rho, antinu, lib=None):
    lib = lib or oracle_lib()
    fn = lib.orc_matter_masses if lib is oracle_lib() else lib.ref_matter_masses
    mm, mv = np.empty(9), np.empty(9)
    fn(_d(U18), _d(DM9), ctypes.c_double(E), ctypes.c_double(rho), ctypes.c_int(int(antinu)), _d_out(mm), _d_out(mv))
    return mm.reshape(3, 3), mv.reshape(3, 3)


def transition_matrix(U18, DM9, antinu, E, rho, Lkm, lib=None):
    """One layer's flavour-basis amplitude matrix A[n, m] (complex128)."""
    lib = lib or oracle_lib()
    fn = lib.orc_transition_matrix if lib is oracle_lib() else lib.ref_transition_matrix
    A = np.empty(18)
    fn(_d(U18), _d(DM9), ctypes.c_int(int(antinu)), ctypes.c_double(E), ctypes.c_double(rho),
       ctypes.c_double(Lkm), _d_out(A))
    return (A[0::2] + 1j * A[1::2]).reshape(3, 3)


def layer_distances(radii_desc, max_layer, cz, h_cm):
    out = np.empty(max_layer + 1)
    oracle_lib().orc_layer_distances(_d(radii_desc), ctypes.c_int(int(max_layer)), ctypes.c_double(cz),
                                     ctypes.c_double(h_cm), _d_out(out))
    return out


def calculate(antinu, cosines, energies, radii_desc, rhos_desc, maxlayers_, h_cm, U18, DM9, threads=0):
    """physics::calculate restated, on already set-up inputs.  Returns P[ic, ie, t]."""
    cosines = np.ascontiguousarray(cosines, dtype=np.float64)
    energies = np.ascontiguousarray(energies, dtype=np.float64)
    ml = np.ascontiguousarray(maxlayers_, dtype=np.int32)
    out = np.empty((len(cosines), len(energies), 9))
    oracle_lib().orc_calculate(ctypes.c_int(int(antinu)), _d(cosines), ctypes.c_int(len(cosines)), _d(energies),
                               ctypes.c_int(len(energies)), _d(radii_desc), _d(rhos_desc), ml.ctypes.data_as(_ip),
                               ctypes.c_double(h_cm), _d(U18), _d(DM9), ctypes.c_int(threads), _d_out(out))
    return out
