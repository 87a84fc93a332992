"""Batched parameter scans on one MI355X: many hypotheses, each with its own uneven grid and
Earth model, propagated in one persistent launch (mixed precision, float results).

Each set's host set-up (PMNS matrix, mass differences, density flip, maxlayers, height)
is the same ``Propagator`` set-up as for a single propagator, so a set's results equal what
``HipPropagatorSingle(..., precision="mixed")`` gives for it.  A scan step changes hypotheses
with :meth:`BatchPropagator.set_params` and calls :meth:`compute`; with ``graph=True`` the
step is a replay of a captured HIP graph.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from .propagator import Propagator


@dataclass
class SetSpec:
    cosines: np.ndarray
    energies: np.ndarray
    radii: np.ndarray          # density-file order, either direction
    rhos: np.ndarray
    params: tuple              # theta12, theta13, theta23, dcp, dm12sq, dm23sq
    antinu: int = 0
    h_km: float = 22.0


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def host_setup(spec: SetSpec) -> Propagator:
    """The single-propagator host set-up of one set (propagator.hpp:117-307)."""
    p = Propagator(len(spec.cosines), len(spec.energies))
    p.setEnergyList(spec.energies)
    p.setCosineList(spec.cosines)
    p.setDensity(spec.radii, spec.rhos)
    p.setProductionHeight(spec.h_km)
    p.setMNSMatrix(*spec.params[:4])
    p.setNeutrinoMasses(*spec.params[4:])
    return p


def mixing_arrays(params: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """U (n, 18) and DM (n, 9) of many hypotheses at once, with the formulas of
    propagator.hpp:191-248 applied elementwise (numpy's sin/cos may differ from the C
    library's in the last bit; use set_params for the per-set C-library path)."""
    params = np.asarray(params, np.float64).reshape(-1, 6)
    t12, t13, t23, d, dm12, dm23 = params.T
    s12, s13, s23 = np.sin(t12), np.sin(t13), np.sin(t23)
    c12, c13, c23 = np.cos(t12), np.cos(t13), np.cos(t23)
    sd, cd = np.sin(d), np.cos(d)
    U = np.zeros((len(params), 18))
    U[:, 0] = c12 * c13
    U[:, 2] = s12 * c13
    U[:, 4], U[:, 5] = s13 * cd, -s13 * sd
    U[:, 6], U[:, 7] = -s12 * c23 - c12 * s23 * s13 * cd, -c12 * s23 * s13 * sd
    U[:, 8], U[:, 9] = c12 * c23 - s12 * s23 * s13 * cd, -s12 * s23 * s13 * sd
    U[:, 10] = s23 * c13
    U[:, 12], U[:, 13] = s12 * s23 - c12 * c23 * s13 * cd, -c12 * c23 * s13 * sd
    U[:, 14], U[:, 15] = -c12 * s23 - s12 * c23 * s13 * cd, -s12 * c23 * s13 * sd
    U[:, 16] = c23 * c13
    m0 = np.where(dm12 == 0.0, -5.0e-9, 0.0)
    m1 = dm12
    m2 = np.where(dm23 == 0.0, (dm12 + dm23) + 5.0e-9, dm12 + dm23)
    D = np.zeros((len(params), 9))
    D[:, 1] = m0 - m1
    D[:, 3] = -D[:, 1]
    D[:, 2] = m0 - m2
    D[:, 6] = -D[:, 2]
    D[:, 5] = m1 - m2
    D[:, 7] = -D[:, 5]
    return U, D


class BatchPropagator:
    def __init__(self, sets: list[SetSpec], device: int = 0, graph: bool = True, autotune: bool = False):
        self._lib = _lib.load()
        self.sets = list(sets)
        self.hosts = [host_setup(s) for s in self.sets]
        n = len(self.sets)
        nc = np.array([h.n_cosines for h in self.hosts], np.int32)
        ne = np.array([h.n_energies for h in self.hosts], np.int32)
        nr = np.array([len(h.radii) for h in self.hosts], np.int32)
        handle = ctypes.c_void_p()
        ip = ctypes.POINTER(ctypes.c_int)
        _lib.check(self._lib.cp3h_batch_create(device, n, nc.ctypes.data_as(ip), ne.ctypes.data_as(ip),
                                               nr.ctypes.data_as(ip), ctypes.byref(handle)))
        self._h = handle
        for i, h in enumerate(self.hosts):
            cos = np.ascontiguousarray(h.cosineList, np.float64)
            ml = np.ascontiguousarray(h.maxlayers, np.int32)
            E = np.ascontiguousarray(h.energyList, np.float64)
            _lib.check(self._lib.cp3h_batch_set_grid_f64(self._h, i, _dp(cos), ml.ctypes.data_as(ip), _dp(E)))
            r = np.ascontiguousarray(h.radii, np.float64)
            d = np.ascontiguousarray(h.rhos, np.float64)
            _lib.check(self._lib.cp3h_batch_set_density_f64(self._h, i, _dp(r), _dp(d)))
            self._upload_params(i)
        self.config = None
        if autotune:
            self.autotune()
        self.use_graph(graph)

    @property
    def n_pairs(self) -> int:
        return sum(h.n_cosines * h.n_energies for h in self.hosts)

    def _upload_params(self, i):
        h, s = self.hosts[i], self.sets[i]
        U = np.ascontiguousarray(h.Mix_U, np.float64)
        D = np.ascontiguousarray(h.dm, np.float64)
        _lib.check(self._lib.cp3h_batch_set_params_f64(self._h, i, _dp(U), _dp(D), int(s.antinu),
                                                       float(h.ProductionHeightinCentimeter)))

    def set_params(self, i: int, params, antinu: int | None = None):
        """New hypothesis for set i (a scan step)."""
        h = self.hosts[i]
        h.setMNSMatrix(*params[:4])
        h.setNeutrinoMasses(*params[4:])
        self.sets[i].params = tuple(params)
        if antinu is not None:
            self.sets[i].antinu = int(antinu)
        self._upload_params(i)

    def set_params_many(self, params, antinu=None):
        """New hypotheses for all sets in one call (params: (n_sets, 6))."""
        params = np.asarray(params, np.float64).reshape(len(self.sets), 6)
        U, D = mixing_arrays(params)
        nu = np.array([s.antinu for s in self.sets] if antinu is None else antinu, np.int32)
        h = np.array([float(x.ProductionHeightinCentimeter) for x in self.hosts])
        for i, s in enumerate(self.sets):
            s.params = tuple(params[i])
            s.antinu = int(nu[i])
        ip = ctypes.POINTER(ctypes.c_int)
        _lib.check(self._lib.cp3h_batch_set_params_many_f64(self._h, 0, len(self.sets), _dp(U), _dp(D),
                                                            nu.ctypes.data_as(ip), _dp(h)))

    def set_binning(self, nodes: int, weights, out_ptr: int, offsets, plane_strides, row_strides):
        """Store only bin averages: set i's (t, ic, bin) average goes to float element
        offsets[i] + t * plane_strides[i] + ic * row_strides[i] + bin of the device buffer at
        out_ptr.  nodes = 0 switches back to unbinned results."""
        n = len(self.sets)
        w = np.zeros(32, np.float32)
        if nodes:
            w[:nodes] = np.asarray(weights, np.float32)
        off = np.ascontiguousarray(np.broadcast_to(np.asarray(offsets if nodes else 0, np.int64), n))
        ps = np.ascontiguousarray(np.broadcast_to(np.asarray(plane_strides if nodes else 0, np.int64), n))
        rs = np.ascontiguousarray(np.broadcast_to(np.asarray(row_strides if nodes else 0, np.int32), n))
        llp = ctypes.POINTER(ctypes.c_longlong)
        _lib.check(self._lib.cp3h_batch_set_binning(self._h, int(nodes), w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                    ctypes.c_void_p(out_ptr or None), off.ctypes.data_as(llp),
                                                    ps.ctypes.data_as(llp), rs.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))

    def set_row_limits(self, i: int, limits):
        """Row ic of set i computes only its first limits[ic] energies (None: all)."""
        if limits is None:
            _lib.check(self._lib.cp3h_batch_set_row_limits(self._h, i, None))
            return
        lim = np.ascontiguousarray(limits, np.int32)
        assert lim.size == len(self.sets[i].cosines)
        _lib.check(self._lib.cp3h_batch_set_row_limits(self._h, i, lim.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))

    def use_graph(self, enable: bool):
        _lib.check(self._lib.cp3h_batch_use_graph(self._h, int(bool(enable))))

    def configure(self, waves_per_group: int, groups_per_cu: int):
        _lib.check(self._lib.cp3h_batch_config(self._h, waves_per_group, groups_per_cu))
        self.config = (waves_per_group, groups_per_cu)

    def autotune(self):
        w, g, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
        _lib.check(self._lib.cp3h_batch_autotune(self._h, ctypes.byref(w), ctypes.byref(g), ctypes.byref(ms)))
        self.config = (w.value, g.value)
        return w.value, g.value, ms.value

    def compute_async(self):
        _lib.check(self._lib.cp3h_batch_compute_async(self._h))

    def wait(self):
        _lib.check(self._lib.cp3h_batch_wait(self._h))

    def compute(self):
        _lib.check(self._lib.cp3h_batch_compute(self._h))

    def last_ms(self) -> tuple[float, float]:
        t, k = ctypes.c_float(), ctypes.c_float()
        _lib.check(self._lib.cp3h_batch_last_ms(self._h, ctypes.byref(t), ctypes.byref(k)))
        return t.value, k.value

    def time_kernels(self, reps: int) -> tuple[float, float]:
        """(set-up, pair) GPU milliseconds per launch, each kernel launched `reps` times back to back
        between two HIP events (cp3h_batch_time_kernels), on the pending hypotheses."""
        su, pa = ctypes.c_float(), ctypes.c_float()
        _lib.check(self._lib.cp3h_batch_time_kernels(self._h, int(reps), ctypes.byref(su), ctypes.byref(pa)))
        return su.value, pa.value

    def results(self, i: int) -> np.ndarray:
        """P of set i as float32 [t, ic, ie] (t = in*3+out)."""
        h = self.hosts[i]
        out = np.empty((9, h.n_cosines, h.n_energies), np.float32)
        _lib.check(self._lib.cp3h_batch_copy_results_f32(self._h, i, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                         out.size))
        return out

    def device_results(self, i: int) -> int:
        """Device address of set i's results: P(t, ic, ie) at [t * plane + ic * row + ie] floats,
        (plane, row) = result_strides(i)."""
        p = ctypes.c_void_p()
        _lib.check(self._lib.cp3h_batch_device_results(self._h, i, ctypes.byref(p)))
        return p.value

    def result_strides(self, i: int) -> tuple[int, int]:
        """(plane, row) strides in floats of set i's device results (rows padded to 32 floats)."""
        plane, row = ctypes.c_longlong(), ctypes.c_int()
        _lib.check(self._lib.cp3h_batch_result_strides(self._h, i, ctypes.byref(plane), ctypes.byref(row)))
        return plane.value, row.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.cp3h_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
