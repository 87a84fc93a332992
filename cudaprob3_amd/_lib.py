"""ctypes binding of libcp3hip.so (C ABI declared in include/cp3hip.h).

The library is built in-tree by ``__graft_entry__.build()``.  Loading fails loudly if it
is missing: there is no CPU fallback for the propagation path.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcp3hip.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "cp3hip.h")

CP3H_OK = 0
CP3H_FP64_PLAIN = 0
CP3H_FP64 = 1
CP3H_MIXED = 2
PRECISIONS = {"fp64_plain": CP3H_FP64_PLAIN, "fp64": CP3H_FP64, "mixed": CP3H_MIXED}
MLP_FORWARD_KERNELS = {"layered": 0, "fused4": 1, "fused8": 2, "fused64": 3}

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_float_p = ctypes.POINTER(ctypes.c_float)
_handle = ctypes.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "cp3h_device_count": (ctypes.c_int, [_c_int_p]),
    "cp3h_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(_handle)]),
    "cp3h_destroy": (ctypes.c_int, [_handle]),
    "cp3h_set_energies_f64": (ctypes.c_int, [_handle, _c_double_p, ctypes.c_int]),
    "cp3h_set_energies_f32": (ctypes.c_int, [_handle, _c_float_p, ctypes.c_int]),
    "cp3h_set_cosines_f64": (ctypes.c_int, [_handle, _c_double_p, _c_int_p, ctypes.c_int]),
    "cp3h_set_cosines_f32": (ctypes.c_int, [_handle, _c_float_p, _c_int_p, ctypes.c_int]),
    "cp3h_set_maxlayers": (ctypes.c_int, [_handle, _c_int_p, ctypes.c_int]),
    "cp3h_set_density_f64": (ctypes.c_int, [_handle, _c_double_p, _c_double_p, ctypes.c_int]),
    "cp3h_set_density_f32": (ctypes.c_int, [_handle, _c_float_p, _c_float_p, ctypes.c_int]),
    "cp3h_set_params_f64": (ctypes.c_int, [_handle, _c_double_p, _c_double_p]),
    "cp3h_set_params_f32": (ctypes.c_int, [_handle, _c_float_p, _c_float_p]),
    "cp3h_compute_async": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_double]),
    "cp3h_wait": (ctypes.c_int, [_handle]),
    "cp3h_compute": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_double]),
    "cp3h_fetch_results": (ctypes.c_int, [_handle, ctypes.POINTER(ctypes.c_void_p), _c_int_p]),
    "cp3h_copy_results_f64": (ctypes.c_int, [_handle, _c_double_p, ctypes.c_size_t]),
    "cp3h_copy_results_f32": (ctypes.c_int, [_handle, _c_float_p, ctypes.c_size_t]),
    "cp3h_device_results": (ctypes.c_int, [_handle, ctypes.POINTER(ctypes.c_void_p), _c_int_p]),
    "cp3h_stream": (ctypes.c_int, [_handle, ctypes.POINTER(ctypes.c_void_p)]),
    "cp3h_set_stream": (ctypes.c_int, [_handle, ctypes.c_void_p]),
    "cp3h_copy_results_device": (ctypes.c_int, [_handle, ctypes.c_void_p, ctypes.c_size_t]),
    "cp3h_sqdiff_results_f32": (ctypes.c_int, [_handle, ctypes.c_void_p, _c_double_p]),
    "cp3h_compute_sqdiff_f32": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_double, ctypes.c_void_p, _c_double_p]),
    "cp3h_last_compute_ms": (ctypes.c_int, [_handle, _c_float_p, _c_float_p]),
    "cp3h_set_timing": (ctypes.c_int, [_handle, ctypes.c_int]),
    "cp3h_time_kernels": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_int,
                                         _c_float_p, _c_float_p]),
    "cp3h_measure_valu_peak": (ctypes.c_int, [ctypes.c_int, _c_float_p, _c_float_p]),
    "cp3h_last_error": (ctypes.c_char_p, []),
    "cp3h_propagate_mixed_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                    ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_int, _c_double_p, _c_double_p, ctypes.c_int,
                                                    ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]),
    "cp3h_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p,
                                         ctypes.POINTER(_handle)]),
    "cp3h_batch_destroy": (ctypes.c_int, [_handle]),
    "cp3h_batch_set_grid_f64": (ctypes.c_int, [_handle, ctypes.c_int, _c_double_p, _c_int_p, _c_double_p]),
    "cp3h_batch_set_density_f64": (ctypes.c_int, [_handle, ctypes.c_int, _c_double_p, _c_double_p]),
    "cp3h_batch_set_params_f64": (ctypes.c_int, [_handle, ctypes.c_int, _c_double_p, _c_double_p, ctypes.c_int,
                                                 ctypes.c_double]),
    "cp3h_batch_set_params_many_f64": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_int, _c_double_p, _c_double_p,
                                                      _c_int_p, _c_double_p]),
    "cp3h_batch_config": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_int]),
    "cp3h_batch_autotune": (ctypes.c_int, [_handle, _c_int_p, _c_int_p, _c_float_p]),
    "cp3h_batch_use_graph": (ctypes.c_int, [_handle, ctypes.c_int]),
    "cp3h_batch_compute_async": (ctypes.c_int, [_handle]),
    "cp3h_batch_wait": (ctypes.c_int, [_handle]),
    "cp3h_batch_compute": (ctypes.c_int, [_handle]),
    "cp3h_batch_last_ms": (ctypes.c_int, [_handle, _c_float_p, _c_float_p]),
    "cp3h_batch_time_kernels": (ctypes.c_int, [_handle, ctypes.c_int, _c_float_p, _c_float_p]),
    "cp3h_batch_copy_results_f32": (ctypes.c_int, [_handle, ctypes.c_int, _c_float_p, ctypes.c_size_t]),
    "cp3h_batch_device_results": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "cp3h_batch_result_strides": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong),
                                                 ctypes.POINTER(ctypes.c_int)]),
    "cp3h_batch_bin_average_f32": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_float_p,
                                                  ctypes.c_void_p]),
    "cp3h_batch_set_binning": (ctypes.c_int, [_handle, ctypes.c_int, _c_float_p, ctypes.c_void_p,
                                              ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                                              _c_int_p]),
    "cp3h_batch_set_row_limits": (ctypes.c_int, [_handle, ctypes.c_int, _c_int_p]),
    "cp3h_mlp_param_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_size_t)]),
    "cp3h_mlp_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_handle)]),
    "cp3h_mlp_destroy": (ctypes.c_int, [_handle]),
    "cp3h_mlp_set_stream": (ctypes.c_int, [_handle, ctypes.c_void_p]),
    "cp3h_mlp_set_scaling": (ctypes.c_int, [_handle, _c_float_p, _c_float_p]),
    "cp3h_mlp_set_params": (ctypes.c_int, [_handle, _c_float_p, ctypes.c_size_t]),
    "cp3h_mlp_get_params": (ctypes.c_int, [_handle, _c_float_p, ctypes.c_size_t]),
    "cp3h_mlp_forward": (ctypes.c_int, [_handle, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p]),
    "cp3h_mlp_set_forward_kernel": (ctypes.c_int, [_handle, ctypes.c_int]),
    "cp3h_mlp_forward_grid": (ctypes.c_int, [_handle, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                             ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "cp3h_mlp_forward_grid_tiles": (ctypes.c_int, [_handle, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                                   ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                   ctypes.c_longlong, ctypes.c_void_p]),
    "cp3h_mlp_gradients": (ctypes.c_int, [_handle, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _c_float_p,
                                          ctypes.c_size_t, _c_float_p]),
    "cp3h_mlp_train_step": (ctypes.c_int, [_handle, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_float,
                                           _c_float_p]),
    "cp3h_mlp_sync": (ctypes.c_int, [_handle]),
}

_LIB: ctypes.CDLL | None = None


class Cp3hError(RuntimeError):
    """A libcp3hip.so call returned a nonzero status."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


def _settle_hip_runtime() -> None:
    """PyTorch wheels bundle their own HIP runtime under the same soname as ROCm's
    (libamdhip64.so.7).  Whichever is loaded first serves the whole process, and PyTorch
    cannot initialise on a foreign one, so when PyTorch is installed it is loaded before
    libcp3hip.so; both then share PyTorch's runtime (one runtime per process)."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() first "
                               "(the propagation path has no CPU fallback)")
        _settle_hip_runtime()
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def check(status: int) -> None:
    if status != CP3H_OK:
        msg = load().cp3h_last_error().decode()
        raise Cp3hError(status, msg)


def device_count() -> int:
    n = ctypes.c_int(0)
    check(load().cp3h_device_count(ctypes.byref(n)))
    return n.value
