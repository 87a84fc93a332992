"""Learned surrogate of energy-bin-averaged oscillation probabilities, with the exact kernel as
its data source, its fallback and the measure of its error.

* Network: 9 -> 256 x 5 (GELU, tanh form) -> 9, BF16 weights, FP32 accumulation, activations
  carried as two bf16 terms; trained and evaluated by the library's matrix-core kernels
  (cp3h_mlp_*; inference runs all layers in one fused launch).  Inputs: cosZ, log10 E, theta12,
  theta13, theta23, deltaCP, dm21^2, dm32^2, nubar flag; outputs the nine P(in -> out).
* Targets: each probability averaged over a +-2.5 % energy bin by 16-point Gauss-Legendre
  quadrature, computed by the exact mixed-precision kernel (batched, cp3h_batch_*) on the
  12-shell jittered Earth.
* Fallback: an error map over (cosZ, log10 E) cells, filled with the largest error seen on
  held-out parameter sets; bins in cells whose bound exceeds the tolerance are recomputed by
  the exact binned kernel.

PyTorch supplies device memory for the datasets and a few tensor reshapes; the arithmetic is
the library's.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from . import workloads as W
from .batch import BatchPropagator, SetSpec

# feature ranges (mapped to [-1, 1]): cosZ, log10 E/GeV, 6 oscillation parameters, flag
LOG10E_RANGE = (math.log10(0.05), 3.0)
FEATURE_LO = np.array([-1.0, LOG10E_RANGE[0], 0.55, 0.14, 0.65, 0.0, 6.8e-5, -2.7e-3, 0.0], np.float32)
FEATURE_HI = np.array([1.0, LOG10E_RANGE[1], 0.62, 0.16, 0.92, 2 * np.pi, 8.0e-5, 2.7e-3, 1.0], np.float32)

GL_NODES, GL_WEIGHTS = np.polynomial.legendre.leggauss(16)

# Earth models a network can be trained for (it is valid for that model only)
EARTHS = {"jitter12": W.jitter12_model, "fine100": W.fine100_model}
BIN_HALF_WIDTH = 0.025


def gl_energies(centers):
    """16 Gauss-Legendre nodes of each +-2.5 % bin, bin-major, and the averaging weights."""
    centers = np.asarray(centers, np.float64)
    E = (centers[:, None] * (1.0 + BIN_HALF_WIDTH * GL_NODES[None, :])).reshape(-1)
    return E, (GL_WEIGHTS / 2.0).astype(np.float32)


def param_count() -> int:
    n = ctypes.c_size_t()
    _lib.check(_lib.load().cp3h_mlp_param_count(ctypes.byref(n)))
    return n.value


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights for the GELU layers, small output layer, zero biases."""
    rng = np.random.Generator(np.random.PCG64(seed))
    shapes = [(256, 9), (256, 256), (256, 256), (256, 256), (256, 256), (9, 256)]
    parts = []
    for i, (n, k) in enumerate(shapes):
        std = math.sqrt(2.0 / k) if i < len(shapes) - 1 else 0.1 / math.sqrt(k)
        parts.append(rng.normal(0.0, std, n * k))
        parts.append(np.zeros(n))
    return np.concatenate(parts).astype(np.float32)


class SurrogateNet:
    def __init__(self, max_batch: int = 8192, training: bool = True, device: int = 0, seed: int = 0):
        import torch
        self.torch = torch
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.cp3h_mlp_create(device, max_batch, int(training), ctypes.byref(h)))
        self._h = h
        self.max_batch = max_batch
        self.device = torch.device(f"cuda:{device}")
        shift = ((FEATURE_LO + FEATURE_HI) / 2).astype(np.float32)
        scale = (2.0 / (FEATURE_HI - FEATURE_LO)).astype(np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        _lib.check(self._lib.cp3h_mlp_set_scaling(self._h, shift.ctypes.data_as(fp), scale.ctypes.data_as(fp)))
        self.shift, self.scale = shift, scale
        # the network runs on a stream torch knows, so the two can be ordered on the device
        self.stream = torch.cuda.Stream(device=self.device)
        self.set_stream(self.stream.cuda_stream)
        self.set_params(init_params(seed))

    def _after_torch(self):
        """Order the network's next launches after torch's pending work (inputs)."""
        self.stream.wait_stream(self.torch.cuda.current_stream(self.device))

    def _before_torch(self):
        """Order torch's next work after the network's launches (outputs)."""
        self.torch.cuda.current_stream(self.device).wait_stream(self.stream)

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        _lib.check(self._lib.cp3h_mlp_set_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.size))

    def get_params(self) -> np.ndarray:
        p = np.empty(param_count(), np.float32)
        _lib.check(self._lib.cp3h_mlp_get_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.size))
        return p

    def set_forward_kernel(self, name: str):
        """'fused4' (default), 'fused8', 'fused64' or 'layered' (the training path's per-layer GEMMs)."""
        _lib.check(self._lib.cp3h_mlp_set_forward_kernel(self._h, _lib.MLP_FORWARD_KERNELS[name]))

    def set_stream(self, stream_ptr):
        _lib.check(self._lib.cp3h_mlp_set_stream(self._h, ctypes.c_void_p(stream_ptr or None)))

    def forward(self, x, out=None):
        """x: cuda float32 [n][9] -> [n][9] (asynchronous on the network's stream)."""
        x = x.contiguous()
        if out is None:
            out = self.torch.empty((x.shape[0], 9), dtype=self.torch.float32, device=x.device)
        self._after_torch()
        _lib.check(self._lib.cp3h_mlp_forward(self._h, ctypes.c_void_p(x.data_ptr()), x.shape[0],
                                              ctypes.c_void_p(out.data_ptr())))
        self._before_torch()
        return out

    def forward_grid(self, cos_t, log10e_t, hyp_t, skip_t, out_t):
        """Grid-mode inference (asynchronous): out_t[h][t][ic][ib] for hypotheses hyp_t [n][7]
        on cos_t x log10e_t; points with ib < skip_t[h][ic] are left to the exact kernel."""
        n_hyp = hyp_t.shape[0]
        assert out_t.numel() >= n_hyp * 9 * cos_t.numel() * log10e_t.numel()
        assert skip_t is None or skip_t.shape[0] >= n_hyp
        self._after_torch()
        _lib.check(self._lib.cp3h_mlp_forward_grid(
            self._h, ctypes.c_void_p(cos_t.data_ptr()), cos_t.numel(), ctypes.c_void_p(log10e_t.data_ptr()),
            log10e_t.numel(), ctypes.c_void_p(hyp_t.data_ptr()), n_hyp,
            ctypes.c_void_p(skip_t.data_ptr() if skip_t is not None else None), ctypes.c_void_p(out_t.data_ptr())))
        self._before_torch()

    def forward_grid_tiles(self, cos_t, log10e_t, hyp_t, skip_t, tiles_t, out_t):
        """Grid-mode inference of the listed 128-row tiles only (tiles_t: device int64 tile
        indices of the h-major row order); other points and those below skip_t stay unwritten."""
        n_hyp = hyp_t.shape[0]
        assert out_t.numel() >= n_hyp * 9 * cos_t.numel() * log10e_t.numel()
        self._after_torch()
        _lib.check(self._lib.cp3h_mlp_forward_grid_tiles(
            self._h, ctypes.c_void_p(cos_t.data_ptr()), cos_t.numel(), ctypes.c_void_p(log10e_t.data_ptr()),
            log10e_t.numel(), ctypes.c_void_p(hyp_t.data_ptr()), n_hyp,
            ctypes.c_void_p(skip_t.data_ptr() if skip_t is not None else None), ctypes.c_void_p(tiles_t.data_ptr()),
            tiles_t.numel(), ctypes.c_void_p(out_t.data_ptr())))
        self._before_torch()

    def sync(self):
        _lib.check(self._lib.cp3h_mlp_sync(self._h))

    def gradients(self, x, t):
        g = np.empty(param_count(), np.float32)
        loss = ctypes.c_float()
        self._after_torch()
        _lib.check(self._lib.cp3h_mlp_gradients(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(t.data_ptr()),
                                                x.shape[0], g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), g.size,
                                                ctypes.byref(loss)))
        return g, loss.value

    def train_step(self, x, t, lr: float, want_loss: bool = False):
        loss = ctypes.c_float()
        self._after_torch()
        _lib.check(self._lib.cp3h_mlp_train_step(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(t.data_ptr()),
                                                 x.shape[0], ctypes.c_float(lr),
                                                 ctypes.byref(loss) if want_loss else None))
        self._before_torch()       # the minibatch tensors may be freed and reused by torch
        return loss.value if want_loss else None

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.cp3h_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --- data from the exact kernel ---------------------------------------------------------------

def features(torch, cos, log10E, params, antinu, device):
    """[n_cos * n_bins][9] rows (cosine-major) for one hypothesis."""
    nc, nb = len(cos), len(log10E)
    X = torch.empty((nc, nb, 9), dtype=torch.float32, device=device)
    X[:, :, 0] = torch.as_tensor(np.asarray(cos, np.float32), device=device)[:, None]
    X[:, :, 1] = torch.as_tensor(np.asarray(log10E, np.float32), device=device)[None, :]
    for k in range(6):
        X[:, :, 2 + k] = float(params[k])
    X[:, :, 8] = float(antinu)
    return X.reshape(nc * nb, 9)


@dataclass
class BinnedSets:
    """Exact bin-averaged probabilities of many hypotheses: X [n][9], T [n][9] on the GPU."""
    X: object
    T: object
    specs: list


def exact_binned(specs_bins, device=0, earth: str = "jitter12"):
    """specs_bins: list of (cosines, bin centres, params, antinu) on the named Earth model.
    Returns per-set device tensors [9][n_cos][n_bins] of GL bin averages."""
    import torch
    radii, rhos = EARTHS[earth]()
    specs, weights = [], None
    for cos, centers, p, nu in specs_bins:
        E, weights = gl_energies(centers)
        specs.append(SetSpec(np.asarray(cos, np.float64), E, radii, rhos, tuple(p), int(nu)))
    b = BatchPropagator(specs, device=device, graph=False)
    b.compute()
    sizes = [9 * len(s.cosines) * (len(s.energies) // 16) for s in specs]
    out = torch.empty(sum(sizes), dtype=torch.float32, device=f"cuda:{device}")
    torch.cuda.current_stream(out.device).synchronize()   # the batch's stream writes into torch memory
    wp = weights.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    _lib.check(b._lib.cp3h_batch_bin_average_f32(b._h, 0, len(specs), 16, wp, ctypes.c_void_p(out.data_ptr())))
    b.close()
    res, off = [], 0
    for (cos, centers, _, _), n in zip(specs_bins, sizes):
        res.append(out[off:off + n].view(9, len(cos), len(centers)))
        off += n
    return res


def random_hypothesis(rng):
    p = list(W.random_params(rng))
    return p, int(rng.random() < 0.5)


def make_dataset(n_sets: int, n_cos: int, n_bins: int, seed: int, device=0, chunk: int = 256,
                 log10e_range=LOG10E_RANGE, earth: str = "jitter12") -> BinnedSets:
    """Random hypotheses (parameter ranges of the scan workload, either ordering, nu or
    nubar), each on n_cos random cosines x n_bins random log-uniform bin centres in
    10**log10e_range GeV (default [0.05, 1000]); exact GL-binned targets."""
    import torch
    rng = np.random.Generator(np.random.PCG64(seed))
    dev = torch.device(f"cuda:{device}")
    Xs, Ts, specs = [], [], []
    for c0 in range(0, n_sets, chunk):
        batch = []
        for _ in range(min(chunk, n_sets - c0)):
            p, nu = random_hypothesis(rng)
            cos = np.sort(rng.uniform(-1.0, 1.0, n_cos))
            l10 = np.sort(rng.uniform(log10e_range[0], log10e_range[1], n_bins))
            batch.append((cos, 10.0 ** l10, p, nu))
        for (cos, centers, p, nu), P in zip(batch, exact_binned(batch, device, earth)):
            Xs.append(features(torch, cos, np.log10(centers), p, nu, dev))
            Ts.append(P.permute(1, 2, 0).reshape(-1, 9))
            specs.append((p, nu))
    return BinnedSets(torch.cat(Xs), torch.cat(Ts), specs)


def train(net: SurrogateNet, data: BinnedSets, epochs: int, batch: int = 8192, lr_max: float = 2e-3,
          lr_min: float = 2e-5, seed: int = 0, log=None):
    """Adam with a cosine learning-rate decay; minibatches drawn without replacement."""
    torch = net.torch
    n = data.X.shape[0]
    steps_per_epoch = n // batch
    total = epochs * steps_per_epoch
    gen = torch.Generator(device=data.X.device)
    gen.manual_seed(seed)
    step = 0
    history = []
    for ep in range(epochs):
        perm = torch.randperm(n, device=data.X.device, generator=gen)
        for s in range(steps_per_epoch):
            idx = perm[s * batch:(s + 1) * batch]
            x, t = data.X[idx].contiguous(), data.T[idx].contiguous()
            lr = lr_min + 0.5 * (lr_max - lr_min) * (1 + math.cos(math.pi * step / max(total, 1)))
            want = s == steps_per_epoch - 1
            loss = net.train_step(x, t, lr, want_loss=want)
            if want:
                history.append(loss)
                if log:
                    log(f"epoch {ep}: loss {loss:.3e} lr {lr:.2e}")
            step += 1
    net.sync()
    return history


# --- error map and fallback ---------------------------------------------------------------------

N_COS_CELLS = 32                 # default error-map cells: 32 cosZ cells x 0.1-decade energy cells,
E_CELL_DECADES = 0.1             # per hypothesis class (nu / nubar x normal / inverted ordering)
N_CLASSES = 4
MIN_CELL_SAMPLES = 32            # a cell seen by fewer held-out samples is never trusted


def n_energy_cells(decades: float) -> int:
    """Energy cells of about `decades` each over LOG10E_RANGE (the last one absorbs the
    remainder, so no cell is a sliver at the range's upper edge)."""
    return max(1, int(round((LOG10E_RANGE[1] - LOG10E_RANGE[0]) / decades)))


def hypothesis_class(params, antinu) -> int:
    return 2 * int(antinu) + int(params[5] < 0)


@dataclass
class ErrorMap:
    """max_err[class][cos cell][energy cell]: the largest held-out |dP| (any channel) in the
    cell, +inf where fewer than MIN_CELL_SAMPLES held-out samples fell.  The cells split
    cosZ in [-1, 1] and log10 E over LOG10E_RANGE evenly (their counts are the array's shape)."""
    max_err: np.ndarray

    def cells(self, cos, log10E):
        n_c, n_e = self.max_err.shape[1], self.max_err.shape[2]
        ic = np.clip(((np.asarray(cos) + 1.0) / 2.0 * n_c).astype(int), 0, n_c - 1)
        width = (LOG10E_RANGE[1] - LOG10E_RANGE[0]) / n_e
        ie = np.clip(((np.asarray(log10E) - LOG10E_RANGE[0]) / width).astype(int), 0, n_e - 1)
        return ic, ie

    def fallback(self, k: int, cos, log10E, tol: float, safety: float) -> np.ndarray:
        """[n_cos][n_bins] mask of the grid points whose cell bound exceeds the tolerance."""
        ic, _ = self.cells(cos, np.zeros(1))
        _, ie = self.cells(np.zeros(1), log10E)
        return self.max_err[k][ic[:, None], ie[None, :]] * safety > tol


def build_error_map(net: SurrogateNet, data: BinnedSets, n_cos_cells: int = N_COS_CELLS,
                    e_cell_decades: float = E_CELL_DECADES, min_samples: int = MIN_CELL_SAMPLES) -> ErrorMap:
    """Largest |dP| of the network on held-out exact data, per class and cell; cells with
    fewer than min_samples held-out samples get +inf (they always fall back)."""
    torch = net.torch
    Y = net.forward(data.X)
    net.sync()
    err = (Y - data.T).abs().amax(dim=1)
    X = data.X
    n_e = n_energy_cells(e_cell_decades)
    width = (LOG10E_RANGE[1] - LOG10E_RANGE[0]) / n_e
    ic = ((X[:, 0] + 1.0) / 2.0 * n_cos_cells).long().clamp(0, n_cos_cells - 1)
    ie = ((X[:, 1] - LOG10E_RANGE[0]) / width).long().clamp(0, n_e - 1)
    k = 2 * X[:, 8].round().long() + (X[:, 7] < 0).long()
    cell = (k * n_cos_cells + ic) * n_e + ie
    n_cells = N_CLASSES * n_cos_cells * n_e
    m = torch.zeros(n_cells, dtype=torch.float32, device=err.device)
    m.scatter_reduce_(0, cell, err, reduce="amax", include_self=False)
    count = torch.bincount(cell, minlength=n_cells)
    m[count < min_samples] = float("inf")
    return ErrorMap(m.view(N_CLASSES, n_cos_cells, n_e).cpu().numpy())


# --- evaluation on a fixed grid ------------------------------------------------------------------

def _hyp_row(params, antinu):
    return [float(x) for x in params] + [float(antinu)]


class GridEvaluator:
    """Bin-averaged probabilities of many hypotheses on one cosZ x energy-bin grid (the jittered
    12-shell Earth): the network fills the grid (fused kernel, grid mode) except, per cosine row,
    the leading run of energy bins up to the last one whose error-map cell exceeds tol / safety;
    the exact mixed-precision kernel computes those bins (batched, binned mode: 16
    Gauss-Legendre nodes per bin, averaged in registers; one set per hypothesis with per-row
    energy limits, so one spectrum table per hypothesis) into the same array, on its own
    stream (the two kernels write disjoint bins).

    Hypotheses are processed `slots` at a time per class; results go to ``self.out``
    [slots][9][n_cos][n_bins] (t = in*3 + out) and are handed to ``consume``.
    """

    def __init__(self, net: SurrogateNet, emap: ErrorMap, cos, centers, tol: float = 1e-3, safety: float = 1.5,
                 slots: int = 8, device: int = 0, earth: str = "jitter12", classes=range(N_CLASSES)):
        torch = net.torch
        self.torch, self.net, self.slots = torch, net, slots
        net.set_forward_kernel("fused4")    # listed tiles run the one-tile-per-workgroup kernel
        self.cos = np.asarray(cos, np.float64)
        self.centers = np.asarray(centers, np.float64)
        n_c, n_b = len(self.cos), len(self.centers)
        self.n_c, self.n_b, self.plane = n_c, n_b, n_c * n_b
        dev = torch.device(f"cuda:{device}")
        self.dev = dev
        l10 = np.log10(self.centers)
        self.cos_t = torch.as_tensor(self.cos.astype(np.float32), device=dev)
        self.l10_t = torch.as_tensor(l10.astype(np.float32), device=dev)
        self.out = torch.empty((slots, 9, n_c, n_b), dtype=torch.float32, device=dev)
        radii, rhos = EARTHS[earth]()
        E_nodes, self.weights = gl_energies(self.centers)
        self.skip, self.skip_t, self.tiles_t, self.batches = [], [], [], []
        for k in range(N_CLASSES):
            if k not in classes:                 # hypotheses of this class are not expected
                self.skip.append(None)
                self.skip_t.append(None)
                self.tiles_t.append(None)
                self.batches.append(None)
                continue
            mask = emap.fallback(k, self.cos, l10, tol, safety)
            last = np.where(mask.any(axis=1), n_b - np.argmax(mask[:, ::-1], axis=1), 0)   # prefix lengths
            self.skip.append(last.astype(np.int32))
            self.skip_t.append(torch.as_tensor(np.tile(last.astype(np.int32), (slots, 1)), device=dev))
            # 128-row tiles (h-major row order over the slots) holding at least one served bin
            served = np.tile((np.arange(n_b)[None, :] >= last[:, None]).reshape(-1), slots)
            pad = (-served.size) % 128
            tiles = np.flatnonzero(np.concatenate([served, np.zeros(pad, bool)]).reshape(-1, 128).any(axis=1))
            self.tiles_t.append(torch.as_tensor(tiles.astype(np.int64), device=dev))
            kmax = int(last.max())
            if kmax == 0:
                self.batches.append(None)
                continue
            dummy = (0.59, 0.15, 0.78, 0.0, 7.4e-5, 2.5e-3 if k % 2 == 0 else -2.5e-3)
            specs = [SetSpec(self.cos, E_nodes[:16 * kmax], radii, rhos, dummy, k // 2) for _ in range(slots)]
            bp = BatchPropagator(specs, device=device, graph=True)
            for i in range(slots):
                bp.set_row_limits(i, 16 * last)
            bp.set_binning(16, self.weights, self.out.data_ptr(), [i * 9 * self.plane for i in range(slots)],
                           self.plane, n_b)
            self.batches.append(bp)
        torch.cuda.synchronize(dev)

    def fallback_fraction(self, k: int) -> float:
        return float(self.skip[k].sum()) / self.plane if self.skip[k] is not None else float("nan")

    def run(self, params, antinu, consume=None) -> dict:
        """Evaluates every hypothesis; returns counts (bins, exact bins) per class."""
        torch, net = self.torch, self.net
        order = sorted(range(len(params)), key=lambda i: hypothesis_class(params[i], antinu[i]))
        stats = {"bins": 0, "exact_bins": 0}
        i = 0
        while i < len(order):
            k = hypothesis_class(params[order[i]], antinu[order[i]])
            if self.skip[k] is None:
                raise ValueError(f"this evaluator was not built for hypothesis class {k}")
            chunk = []
            while i < len(order) and len(chunk) < self.slots and \
                    hypothesis_class(params[order[i]], antinu[order[i]]) == k:
                chunk.append(order[i])
                i += 1
            rows = [_hyp_row(params[j], antinu[j]) for j in chunk]
            rows += [rows[-1]] * (self.slots - len(chunk))        # idle slots repeat the last one
            hyp = torch.tensor(rows, dtype=torch.float32, device=self.dev)
            torch.cuda.synchronize(self.dev)          # previous chunk consumed; hyp uploaded
            net.forward_grid_tiles(self.cos_t, self.l10_t, hyp, self.skip_t[k], self.tiles_t[k], self.out)
            bp = self.batches[k]
            if bp is not None:
                bp.set_params_many(np.array([r[:6] for r in rows], np.float64), np.full(self.slots, k // 2, np.int32))
                bp.compute_async()
            net.sync()
            if bp is not None:
                bp.wait()
            stats["bins"] += len(chunk) * self.plane
            stats["exact_bins"] += len(chunk) * int(self.skip[k].sum())
            if consume is not None:
                consume(chunk, self.out[:len(chunk)])
        return stats

    def close(self):
        for bp in self.batches:
            if bp is not None:
                bp.close()
        self.batches = []


class ExactBinnedGrid:
    """The exact mixed-precision kernel alone on the same grid: every bin from 16 GL nodes
    (batched, binned mode), `slots` hypotheses per launch, into ``self.out``."""

    def __init__(self, torch, cos, centers, slots: int = 8, device: int = 0, earth: str = "jitter12"):
        self.torch, self.slots = torch, slots
        self.cos = np.asarray(cos, np.float64)
        self.centers = np.asarray(centers, np.float64)
        n_c, n_b = len(self.cos), len(self.centers)
        self.plane = n_c * n_b
        self.dev = torch.device(f"cuda:{device}")
        self.out = torch.empty((slots, 9, n_c, n_b), dtype=torch.float32, device=self.dev)
        radii, rhos = EARTHS[earth]()
        E_nodes, w = gl_energies(self.centers)
        specs = [SetSpec(self.cos, E_nodes, radii, rhos, (0.59, 0.15, 0.78, 0.0, 7.4e-5, 2.5e-3), 0)
                 for _ in range(slots)]
        self.bp = BatchPropagator(specs, device=device, graph=True)
        self.bp.set_binning(16, w, self.out.data_ptr(), [s * 9 * self.plane for s in range(slots)], self.plane, n_b)
        torch.cuda.synchronize(self.dev)

    def run(self, params, antinu, consume=None):
        for c0 in range(0, len(params), self.slots):
            idx = list(range(c0, min(c0 + self.slots, len(params))))
            sel = idx + [idx[-1]] * (self.slots - len(idx))
            self.torch.cuda.synchronize(self.dev)
            self.bp.set_params_many(np.array([params[j] for j in sel], np.float64),
                                    np.array([antinu[j] for j in sel], np.int32))
            self.bp.compute()
            if consume is not None:
                consume(idx, self.out[:len(idx)])

    def close(self):
        self.bp.close()
