"""Python mirror of the reference's ``cudaprob3::Propagator<FLOAT_T>`` API and the MI355X
``HipPropagatorSingle`` behind it.

``Propagator`` restates the host-side set-up of propagator.hpp:43-337 with the same
formulas, operation order, precision (``dtype`` = FLOAT_T) and error messages, so that the
arrays handed to the device (U, DM, descending radii/rhos, coslimit, maxlayers, height in
cm) are bit-identical to the reference's.  ``HipPropagatorSingle`` replaces
``CudaPropagatorSingle`` (cudapropagator.cuh:39-251): same constructor forms, same setters,
``calculateProbabilities`` / ``getProbability``, with the kernels in libcp3hip.so.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib

# types.hpp:23-29
e_e, e_m, e_t, m_e, m_m, m_t, t_e, t_m, t_t = range(9)
Neutrino, Antineutrino = 0, 1

R_EARTH_KM = 6371.0          # constants.hpp:34


class Propagator:
    """Host set-up of the reference's abstract ``Propagator`` (propagator.hpp:43-337)."""

    def __init__(self, n_cosines: int, n_energies: int, dtype=np.float64):
        self.n_cosines = int(n_cosines)
        self.n_energies = int(n_energies)
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.dtype(np.float64), np.dtype(np.float32)):
            raise TypeError("FLOAT_T must be float64 or float32")
        self.T = self.dtype.type
        self.energyList = np.zeros(self.n_energies, self.dtype)
        self.cosineList = np.zeros(self.n_cosines, self.dtype)
        self.maxlayers = np.zeros(self.n_cosines, np.int32)
        self.radii = np.zeros(0, self.dtype)
        self.rhos = np.zeros(0, self.dtype)
        self.coslimit = np.zeros(0, self.dtype)
        self.Mix_U = np.zeros(18, self.dtype)   # U(i,j) = (Mix_U[2(3i+j)], Mix_U[2(3i+j)+1])
        self.dm = np.zeros(9, self.dtype)       # DM(i,j) = dm[3i+j]
        self.ProductionHeightinCentimeter = self.T(0)
        self.isSetProductionHeight = False
        self.isSetCosine = False
        self.isInit = True

    # ---- arithmetic in FLOAT_T ------------------------------------------------------
    def _libm(self, fn, x):
        # the reference calls the global double overloads of sin/cos/sqrt (propagator.hpp:155,
        # 193-201) and stores the result as FLOAT_T
        return self.T(fn(float(x)))

    # propagator.hpp:117-161
    def setDensity(self, radii_, rhos_):
        radii_ = np.asarray(radii_, dtype=self.dtype)
        rhos_ = np.asarray(rhos_, dtype=self.dtype)
        if len(rhos_) != len(radii_):
            raise RuntimeError("setDensity : rhos.size() != radii.size()")
        if len(rhos_) == 0 or len(radii_) == 0:
            raise RuntimeError("setDensity : vectors must not be empty")
        need_flip = False
        if len(radii_) >= 2:
            sign = 1 if radii_[1] - radii_[0] > 0 else -1
            for i in range(1, len(radii_)):
                if (radii_[i] - radii_[i - 1]) * sign < 0:
                    raise RuntimeError("radii order messed up")
            need_flip = sign == 1
        self.radii = radii_.copy()
        self.rhos = rhos_.copy()
        if need_flip:
            self.radii = self.radii[::-1].copy()
            self.rhos = self.rhos[::-1].copy()
        T = self.T
        R = T(R_EARTH_KM)
        lim = []
        for i, r in enumerate(self.radii):
            q = T(1) - (T(r) * T(r) / (R * R))
            x = T(-1 * math.sqrt(float(q))) if float(q) >= 0 else T(float("nan"))
            if i == 0:
                x = T(0)
            lim.append(x)
        self.coslimit = np.array(lim, dtype=self.dtype)
        self.setMaxlayers()

    # propagator.hpp:169-184
    def setDensityFromFile(self, filename: str):
        try:
            with open(filename) as f:
                text = f.read()
        except OSError:
            raise RuntimeError("could not open density file " + filename) from None
        radii, rhos = [], []
        tokens = text.split()
        for k in range(0, len(tokens) - 1, 2):     # stops at the first unreadable pair
            try:
                r, d = float(tokens[k]), float(tokens[k + 1])
            except ValueError:
                break
            radii.append(r)
            rhos.append(d)
        self.setDensity(np.array(radii, dtype=self.dtype), np.array(rhos, dtype=self.dtype))

    # propagator.hpp:191-221
    def setMNSMatrix(self, theta12, theta13, theta23, dCP):
        T = self.T
        s12, s13, s23 = (self._libm(math.sin, T(t)) for t in (theta12, theta13, theta23))
        c12, c13, c23 = (self._libm(math.cos, T(t)) for t in (theta12, theta13, theta23))
        sd = self._libm(math.sin, T(dCP))
        cd = self._libm(math.cos, T(dCP))
        U = np.zeros(18, self.dtype)

        def put(i, j, re, im):
            U[2 * (3 * i + j)] = re
            U[2 * (3 * i + j) + 1] = im

        z = T(0)
        put(0, 0, c12 * c13, z)
        put(0, 1, s12 * c13, z)
        put(0, 2, s13 * cd, -s13 * sd)
        put(1, 0, -s12 * c23 - c12 * s23 * s13 * cd, -c12 * s23 * s13 * sd)
        put(1, 1, c12 * c23 - s12 * s23 * s13 * cd, -s12 * s23 * s13 * sd)
        put(1, 2, s23 * c13, z)
        put(2, 0, s12 * s23 - c12 * c23 * s13 * cd, -c12 * c23 * s13 * sd)
        put(2, 1, -c12 * s23 - s12 * c23 * s13 * cd, -s12 * c23 * s13 * sd)
        put(2, 2, c23 * c13, z)
        self.Mix_U = U

    # propagator.hpp:226-248
    def setNeutrinoMasses(self, dm12sq, dm23sq):
        T = self.T
        dm12sq, dm23sq = T(dm12sq), T(dm23sq)
        mv = [T(0.0), dm12sq, dm12sq + dm23sq]
        delta = T(5.0e-9)
        if dm12sq == 0.0:
            mv[0] = mv[0] - delta
        if dm23sq == 0.0:
            mv[2] = mv[2] + delta
        D = np.zeros(9, self.dtype)
        D[1] = mv[0] - mv[1]
        D[3] = -D[1]
        D[2] = mv[0] - mv[2]
        D[6] = -D[2]
        D[5] = mv[1] - mv[2]
        D[7] = -D[5]
        self.dm = D

    # propagator.hpp:252-257
    def setEnergyList(self, lst):
        lst = np.asarray(lst, dtype=self.dtype)
        if len(lst) != self.n_energies:
            raise RuntimeError("Propagator::setEnergyList. Propagator was not created for this number of energy nodes")
        self.energyList = lst.copy()

    # propagator.hpp:261-274
    def setCosineList(self, lst):
        lst = np.asarray(lst, dtype=self.dtype)
        if len(lst) != self.n_cosines:
            raise RuntimeError("Propagator::setCosineList. Propagator was not created for this number of cosine nodes")
        self.cosineList = lst.copy()
        if self.isSetProductionHeight:
            self.setProductionHeight(self.T(self.ProductionHeightinCentimeter / 100000.0))
        self.setMaxlayers()
        self.isSetCosine = True

    # propagator.hpp:279-286
    def setProductionHeight(self, heightKM):
        if not self.isSetCosine:
            raise RuntimeError("must set cosine list before production height")
        self.ProductionHeightinCentimeter = self.T(self.T(heightKM) * 100000.0)
        self.isSetProductionHeight = True

    # propagator.hpp:301-307
    def setMaxlayers(self):
        lim = self.coslimit
        if len(lim) == 0:
            self.maxlayers = np.zeros(self.n_cosines, np.int32)
        else:
            self.maxlayers = (self.cosineList[:, None] < lim[None, :]).sum(axis=1).astype(np.int32)

    def calculateProbabilities(self, type_):
        raise NotImplementedError

    def getProbability(self, index_cosine, index_energy, t):
        raise NotImplementedError


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class HipPropagatorSingle(Propagator):
    """Single-GPU propagator on MI355X; drop-in for ``CudaPropagatorSingle``.

    ``HipPropagatorSingle(id, n_cosines, n_energies)`` or ``(n_cosines, n_energies)``
    (cudapropagator.cuh:50, 86).  ``precision``: ``"fp64_plain"`` (the reference's per-pair
    FP64 algorithm), ``"fp64"`` (hoisted FP64 spectrum, same parity bar; default for float64,
    as in the C++ class), ``"mixed"`` (FP32 bulk with FP64 eigenvalues and phases, float
    results; default for float32).
    """

    def __init__(self, *args, dtype=np.float64, precision: str | None = None):
        if len(args) == 3:
            device, n_c, n_e = args
        elif len(args) == 2:
            device, (n_c, n_e) = 0, args
        else:
            raise TypeError("HipPropagatorSingle(id, n_cosines, n_energies) or (n_cosines, n_energies)")
        super().__init__(n_c, n_e, dtype)
        if precision is None:
            precision = "fp64" if self.dtype == np.float64 else "mixed"
        if precision not in _lib.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_lib.PRECISIONS)}")
        self.precision = precision
        self.deviceId = int(device)
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.cp3h_create(self.deviceId, self.n_cosines, self.n_energies,
                                         _lib.PRECISIONS[precision], ctypes.byref(h)))
        self._h = h
        self._host_result = None
        self.resultsResideOnHost = False

    # ---- lifetime -----------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.cp3h_destroy(self._h)
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

    def _handle(self):
        if getattr(self, "_h", None) is None:
            raise RuntimeError("HipPropagatorSingle::calculateProbabilities. Object has been moved from.")
        return self._h

    # ---- setters forwarding to the device (cudapropagator.cuh:132-188) -------------------
    def setDensity(self, radii_, rhos_):
        super().setDensity(radii_, rhos_)
        r = np.ascontiguousarray(self.radii, dtype=np.float64)
        d = np.ascontiguousarray(self.rhos, dtype=np.float64)
        _lib.check(self._lib.cp3h_set_density_f64(self._handle(), _ptr(r, ctypes.c_double),
                                                  _ptr(d, ctypes.c_double), len(r)))

    def setEnergyList(self, lst):
        super().setEnergyList(lst)
        e = np.ascontiguousarray(self.energyList, dtype=np.float64)
        _lib.check(self._lib.cp3h_set_energies_f64(self._handle(), _ptr(e, ctypes.c_double), len(e)))

    def setCosineList(self, lst):
        super().setCosineList(lst)
        c = np.ascontiguousarray(self.cosineList, dtype=np.float64)
        m = np.ascontiguousarray(self.maxlayers, dtype=np.int32)
        _lib.check(self._lib.cp3h_set_cosines_f64(self._handle(), _ptr(c, ctypes.c_double),
                                                  _ptr(m, ctypes.c_int), len(c)))

    def setMaxlayers(self):
        super().setMaxlayers()
        if getattr(self, "_h", None) is not None:
            m = np.ascontiguousarray(self.maxlayers, dtype=np.int32)
            _lib.check(self._lib.cp3h_set_maxlayers(self._h, _ptr(m, ctypes.c_int), len(m)))

    # ---- compute (cudapropagator.cuh:162-226) -------------------------------------------
    def calculateProbabilitiesAsync(self, type_):
        """Like the reference (cudapropagator.cuh:191-220), every computation hands the current
        mixing matrix and mass differences over again (the library then redoes their host
        set-up: vacuum root order and the mixed kernels' real basis)."""
        h = self._handle()
        if not self.isSetProductionHeight:
            raise RuntimeError("HipPropagatorSingle::calculateProbabilities. production height was not set")
        self.resultsResideOnHost = False
        self._upload_params(h)
        _lib.check(self._lib.cp3h_compute_async(h, int(type_), float(self.ProductionHeightinCentimeter)))

    def computeSqdiff(self, type_, data_ptr: int) -> float:
        """calculateProbabilities fused with sum (P - data)^2 over all results (FP64, on the
        device; mixed precision only).  data_ptr: device float array in the results' SoA
        layout.  The stored results are not updated."""
        h = self._handle()
        if not self.isSetProductionHeight:
            raise RuntimeError("HipPropagatorSingle::calculateProbabilities. production height was not set")
        self._upload_params(h)
        r = ctypes.c_double()
        _lib.check(self._lib.cp3h_compute_sqdiff_f32(h, int(type_), float(self.ProductionHeightinCentimeter),
                                                     ctypes.c_void_p(data_ptr), ctypes.byref(r)))
        return r.value

    def _upload_params(self, h):
        """Hands U and DM to the handle (cp3h_set_params_*, which redoes the host set-up derived
        from them).  The marshalled arrays are kept while Mix_U and dm are the same objects."""
        key = (id(self.Mix_U), id(self.dm))
        cached = getattr(self, "_param_args", None)
        if cached is None or cached[0] != key:
            ctype = ctypes.c_double if self.dtype == np.float64 else ctypes.c_float
            U = np.ascontiguousarray(self.Mix_U, dtype=self.dtype)
            D = np.ascontiguousarray(self.dm, dtype=self.dtype)
            fn = self._lib.cp3h_set_params_f64 if self.dtype == np.float64 else self._lib.cp3h_set_params_f32
            cached = (key, fn, _ptr(U, ctype), _ptr(D, ctype), U, D, self.Mix_U, self.dm)
            self._param_args = cached
        _lib.check(cached[1](h, cached[2], cached[3]))

    def waitForCompletion(self):
        _lib.check(self._lib.cp3h_wait(self._handle()))

    def calculateProbabilities(self, type_):
        self.calculateProbabilitiesAsync(type_)
        self.waitForCompletion()

    def getResultFromDevice(self):
        """cudapropagator.cuh:229-235: D2H into the handle's pinned host buffer (cp3h_fetch_results),
        then a copy in FLOAT_T (the pinned buffer is overwritten by the next fetch)."""
        n = 9 * self.n_cosines * self.n_energies
        ptr, nbytes = ctypes.c_void_p(), ctypes.c_int()
        _lib.check(self._lib.cp3h_fetch_results(self._handle(), ctypes.byref(ptr), ctypes.byref(nbytes)))
        ctype = ctypes.c_double if nbytes.value == 8 else ctypes.c_float
        if n:
            pinned = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), shape=(n,))
            out = pinned.astype(self.dtype, copy=True)
        else:
            out = np.empty(0, dtype=self.dtype)
        self._host_result = out.reshape(9, self.n_cosines, self.n_energies)
        self.resultsResideOnHost = True

    # cudapropagator.cuh:168-181
    def getProbability(self, index_cosine, index_energy, t):
        if index_cosine >= self.n_cosines or index_energy >= self.n_energies:
            raise RuntimeError("HipPropagatorSingle::getProbability. Invalid indices")
        if index_cosine < 0 or index_energy < 0 or not 0 <= int(t) < 9:
            raise RuntimeError("HipPropagatorSingle::getProbability. Invalid indices")
        if not self.resultsResideOnHost:
            self.getResultFromDevice()
        return self._host_result[int(t), index_cosine, index_energy]

    # ---- extensions ---------------------------------------------------------------------
    def getProbabilities(self) -> np.ndarray:
        """All results as an array [t, ic, ie] (SoA, t = in*3+out)."""
        if not self.resultsResideOnHost:
            self.getResultFromDevice()
        return self._host_result

    def device_results(self):
        """(device pointer, element bytes) of the SoA result buffer."""
        p, b = ctypes.c_void_p(), ctypes.c_int()
        _lib.check(self._lib.cp3h_device_results(self._handle(), ctypes.byref(p), ctypes.byref(b)))
        return p.value, b.value

    def set_stream(self, stream_ptr: int | None):
        """Enqueue on the caller's hipStream_t (an int handle, e.g. torch.cuda.Stream().cuda_stream);
        None returns to the propagator's own stream."""
        _lib.check(self._lib.cp3h_set_stream(self._handle(), ctypes.c_void_p(stream_ptr or None)))

    def set_timing(self, enable: bool):
        """Record HIP events around every compute (for last_compute_times); off by default."""
        _lib.check(self._lib.cp3h_set_timing(self._handle(), int(bool(enable))))

    def time_kernels(self, type_, reps: int, data_ptr: int | None = None) -> tuple[float, float]:
        """(set-up, per-pair) GPU milliseconds per launch, each kernel launched `reps` times back to
        back between two HIP events (cp3h_time_kernels); data_ptr times the likelihood variant."""
        h = self._handle()
        if not self.isSetProductionHeight:
            raise RuntimeError("HipPropagatorSingle::calculateProbabilities. production height was not set")
        self._upload_params(h)
        self.resultsResideOnHost = False
        su, pa = ctypes.c_float(), ctypes.c_float()
        _lib.check(self._lib.cp3h_time_kernels(h, int(type_), float(self.ProductionHeightinCentimeter),
                                               ctypes.c_void_p(data_ptr or None), int(reps), ctypes.byref(su),
                                               ctypes.byref(pa)))
        return su.value, pa.value

    def last_compute_ms(self) -> float:
        """GPU milliseconds of all kernels of the last compute."""
        return self.last_compute_times()[0]

    def last_compute_times(self) -> tuple[float, float]:
        """(all kernels, per-pair kernel only) GPU milliseconds of the last compute."""
        total, pairs = ctypes.c_float(), ctypes.c_float()
        _lib.check(self._lib.cp3h_last_compute_ms(self._handle(), ctypes.byref(total), ctypes.byref(pairs)))
        return total.value, pairs.value
