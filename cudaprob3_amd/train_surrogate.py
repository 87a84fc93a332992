"""Trains the surrogate network on exact bin-averaged probabilities and writes its parameters
and error map (cudaprob3_amd/data/surrogate_<earth>.npz).

    python -m cudaprob3_amd.train_surrogate [--earth jitter12|fine100] [--samples 50331648]
                                            [--epochs 80] [--lr 1e-3] [--min-energy 2]

Training data: `samples` points (cosZ, energy-bin centre, hypothesis) in sets of 64 random
cosines x 64 log-uniform bin centres per random hypothesis (parameter ranges of the scan
workload, either ordering, nu or nubar), targets from the exact mixed-precision kernel
averaged over +-2.5 % bins by 16-point Gauss-Legendre quadrature, on one Earth model (the
jittered 12-shell profile, or the 100-shell profile of the Markov-chain workload).  Bins below --min-energy GeV are left out of training (they always fall back to the
exact kernel).  The error map comes from 1024 further hypotheses over the full energy range.
Runs on one GPU in a few minutes.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

import numpy as np

from . import surrogate as S

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_PATH = os.path.join(DATA_DIR, "surrogate_jitter12.npz")


def path_for(earth: str) -> str:
    return os.path.join(DATA_DIR, f"surrogate_{earth}.npz")


def save(path, net, emap: S.ErrorMap, meta: dict):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez_compressed(path, params=net.get_params(), max_err=emap.max_err.astype(np.float32),
                        feature_lo=S.FEATURE_LO, feature_hi=S.FEATURE_HI, meta=np.array(json.dumps(meta)))


def load(path: str = DEFAULT_PATH, max_batch: int = 8192, device: int = 0):
    """(network for inference, error map, metadata) from a file written by save()."""
    with np.load(path, allow_pickle=False) as z:
        if not (np.array_equal(z["feature_lo"], S.FEATURE_LO) and np.array_equal(z["feature_hi"], S.FEATURE_HI)):
            raise ValueError(f"{path}: input scaling differs from this build's")
        net = S.SurrogateNet(max_batch=max_batch, training=False, device=device)
        net.set_params(z["params"])
        return net, S.ErrorMap(z["max_err"].astype(np.float64)), json.loads(str(z["meta"]))


def rebuild_map(path: str, held_sets: int, n_cos_cells: int, e_cell_decades: float, min_samples: int,
                seed: int | None = None):
    """Recomputes the error map of a saved network on `held_sets` fresh held-out hypotheses (same
    Earth model; seed default: the training seed + 1) and rewrites the file."""
    net, _, meta = load(path, max_batch=65536)
    earth = meta.get("earth", "jitter12")
    seed = meta["seed"] + 1 if seed is None else seed
    held = S.make_dataset(n_sets=held_sets, n_cos=64, n_bins=64, seed=seed, earth=earth)
    emap = S.build_error_map(net, held, n_cos_cells=n_cos_cells, e_cell_decades=e_cell_decades,
                             min_samples=min_samples)
    meta.update(earth=earth, held_out_samples=held.X.shape[0], held_out_seed=seed,
                error_map={"cos_cells": n_cos_cells, "energy_cells": int(emap.max_err.shape[2]),
                           "min_cell_samples": min_samples})
    save(path, net, emap, meta)
    net.close()
    return emap, meta


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--samples", type=int, default=50_331_648)
    ap.add_argument("--epochs", type=int, default=80)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--min-energy", type=float, default=2.0)
    ap.add_argument("--seed", type=int, default=101)
    ap.add_argument("--earth", default="jitter12", choices=sorted(S.EARTHS))
    ap.add_argument("--out", default=None, help="default: cudaprob3_amd/data/surrogate_<earth>.npz")
    ap.add_argument("--held-sets", type=int, default=1024, help="held-out hypotheses of the error map")
    ap.add_argument("--cos-cells", type=int, default=S.N_COS_CELLS)
    ap.add_argument("--e-cell-decades", type=float, default=S.E_CELL_DECADES)
    ap.add_argument("--min-cell-samples", type=int, default=S.MIN_CELL_SAMPLES)
    ap.add_argument("--rebuild-map", action="store_true",
                    help="keep the saved network, recompute only its error map")
    a = ap.parse_args(argv)
    out = a.out or path_for(a.earth)
    if a.rebuild_map:
        _, meta = rebuild_map(out, a.held_sets, a.cos_cells, a.e_cell_decades, a.min_cell_samples)
        print(json.dumps(meta), flush=True)
        return
    t0 = time.time()
    lo = math.log10(max(a.min_energy, 10 ** S.LOG10E_RANGE[0]))
    n_sets = a.samples // 4096
    data = S.make_dataset(n_sets=n_sets, n_cos=64, n_bins=64, seed=a.seed, log10e_range=(lo, S.LOG10E_RANGE[1]),
                          earth=a.earth)
    print(f"training data: {data.X.shape[0]} samples from {n_sets} hypotheses ({time.time() - t0:.1f} s)", flush=True)
    net = S.SurrogateNet(max_batch=65536, training=True, seed=3)
    t1 = time.time()
    hist = S.train(net, data, epochs=a.epochs, batch=a.batch, lr_max=a.lr, seed=a.seed,
                   log=lambda s: print(f"[{time.time() - t1:.0f} s] {s}", flush=True))
    del data
    held = S.make_dataset(n_sets=a.held_sets, n_cos=64, n_bins=64, seed=a.seed + 1, earth=a.earth)
    emap = S.build_error_map(net, held, n_cos_cells=a.cos_cells, e_cell_decades=a.e_cell_decades,
                             min_samples=a.min_cell_samples)
    meta = dict(earth=a.earth, samples=n_sets * 4096, epochs=a.epochs, batch=a.batch, lr=a.lr,
                min_energy_gev=10 ** lo, seed=a.seed,
                final_loss=hist[-1] if hist else None, train_seconds=time.time() - t1, held_out_samples=held.X.shape[0],
                held_out_seed=a.seed + 1,
                error_map={"cos_cells": a.cos_cells, "energy_cells": int(emap.max_err.shape[2]),
                           "min_cell_samples": a.min_cell_samples})
    save(out, net, emap, meta)
    print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main()
