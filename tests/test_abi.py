"""The C ABI library loads and exports exactly what include/cp3hip.h declares (CPU only:
no call here reaches a GPU; only host-side argument checks run)."""
import ctypes
import os
import re
import subprocess

import pytest

from cudaprob3_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(_lib.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cp3h_\w+)\s*\(", text)))


def test_library_built_and_loads():
    assert os.path.exists(_lib.LIB_PATH), "run __graft_entry__.build()"
    _lib.load()


def test_every_declared_symbol_is_exported_and_bound():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), f"{n} declared in cp3hip.h but not exported"
    assert sorted(_lib.SIGNATURES) == names


def test_null_handle_is_rejected_without_device_work():
    lib = _lib.load()
    assert lib.cp3h_destroy(None) == 0
    rc = lib.cp3h_set_energies_f64(None, None, 0)
    assert rc == 1 and "null propagator handle" in lib.cp3h_last_error().decode()
    rc = lib.cp3h_compute(None, 0, 0.0)
    assert rc == 1
    out = ctypes.c_void_p()
    rc = lib.cp3h_create(0, 4, 4, 7, ctypes.byref(out))
    assert rc == 1 and "unknown precision mode" in lib.cp3h_last_error().decode()
    su, pa = ctypes.c_float(), ctypes.c_float()
    rc = lib.cp3h_time_kernels(None, 0, 0.0, None, 5, ctypes.byref(su), ctypes.byref(pa))
    assert rc == 1 and "null propagator handle" in lib.cp3h_last_error().decode()
    ptr, nbytes = ctypes.c_void_p(), ctypes.c_int()
    rc = lib.cp3h_fetch_results(None, ctypes.byref(ptr), ctypes.byref(nbytes))
    assert rc == 1 and "null propagator handle" in lib.cp3h_last_error().decode()
    rc = lib.cp3h_batch_time_kernels(None, 5, ctypes.byref(su), ctypes.byref(pa))
    assert rc == 1 and "null batch handle" in lib.cp3h_last_error().decode()
    rc = lib.cp3h_measure_valu_peak(0, None, None)
    assert rc == 1 and "null output pointers" in lib.cp3h_last_error().decode()


def test_cpp_headers_compile():
    """The drop-in C++ API (Propagator<FLOAT_T>, HipPropagatorSingle<FLOAT_T>) compiles for
    both FLOAT_T instantiations with the host compiler."""
    src = os.path.join(ROOT, "tests", "cpp", "dropin.cpp")
    if not os.path.exists(src):
        pytest.skip("C++ drop-in driver not present")
    r = subprocess.run(["g++", "-std=c++14", "-fsyntax-only", "-Wall", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
