"""Write the golden fixtures of the propagation path from the reference itself.

Runs the reference CpuPropagator<double>, compiled in place from the reference headers by
``make -C oracle ref`` (oracle/_ref/libcp3ref.so), on small grids and stores the exact
inputs (float64) with the outputs.  The reference ships no expected outputs of its own
(no tests, no golden vectors), so these files are the reference-held pins of the path.

    python tests/golden/make_golden.py        # writes tests/golden/*.npz

Each fixture: cosines, energies, radii, rhos (density-file order), params
(theta12, theta13, theta23, dcp, dm12sq, dm23sq), h_km, antinu, P[ic, ie, t] (AoS).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from cudaprob3_amd import workloads as W  # noqa: E402


def cases():
    g = W.prem4_256()
    for antinu in (0, 1):
        yield f"prem4_256_strided_{'nubar' if antinu else 'nu'}", dict(
            cosines=g.cosines[::8], energies=g.energies[::8], radii=g.radii, rhos=g.rhos,
            params=g.params, h_km=g.h_km, antinu=antinu)
    cos16 = np.linspace(-1.0, 1.0, 16)
    e16 = np.geomspace(0.1, 100.0, 16)
    io = (0.5695951908800630, 0.1608752771983211, 0.7853981633974483, 1.2, 7.9e-5, -2.5e-3)
    yield "prem12_io_dcp", dict(cosines=cos16, energies=e16, radii=W.PREM12[:, 0], rhos=W.PREM12[:, 1],
                                params=io, h_km=22.0, antinu=0)
    yield "prem12_io_dcp_nubar", dict(cosines=cos16, energies=e16, radii=W.PREM12[:, 0], rhos=W.PREM12[:, 1],
                                      params=io, h_km=22.0, antinu=1)
    yield "prem59_nu", dict(cosines=cos16, energies=e16, radii=W.PREM59[:, 0], rhos=W.PREM59[:, 1],
                            params=W.NUFIT, h_km=15.0, antinu=0)
    # descending density-file order, extreme cosines, tiny and huge energies
    edge_cos = np.array([-1.0, -0.9999999, -0.83, -0.5, -1e-9, 0.0, 1e-9, 0.4, 1.0])
    edge_e = np.array([0.05, 0.1, 0.7, 3.0, 25.0, 1000.0])
    yield "edge_prem4_desc", dict(cosines=edge_cos, energies=edge_e, radii=W.PREM4[::-1, 0].copy(),
                                  rhos=W.PREM4[::-1, 1].copy(), params=W.NUFIT, h_km=22.0, antinu=0)
    # degeneracy nudges of propagator.hpp:233-236
    yield "degenerate_dm12_zero", dict(cosines=edge_cos, energies=edge_e, radii=W.PREM4[:, 0], rhos=W.PREM4[:, 1],
                                       params=(0.58, 0.15, 0.72, 0.0, 0.0, 2.5e-3), h_km=22.0, antinu=0)
    yield "degenerate_dm23_zero", dict(cosines=edge_cos, energies=edge_e, radii=W.PREM4[:, 0], rhos=W.PREM4[:, 1],
                                       params=(0.58, 0.15, 0.72, 0.5, 7.4e-5, 0.0), h_km=22.0, antinu=1)
    # single row model (the surface only): every path is vacuum
    yield "single_row_model", dict(cosines=edge_cos, energies=edge_e, radii=np.array([6371.0]),
                                   rhos=np.array([3.3]), params=W.NUFIT, h_km=22.0, antinu=0)
    for g in W.jitter12_variants(2048):
        tag = f"{'nubar' if g.antinu else 'nu'}_{'io' if g.params[5] < 0 else 'no'}"
        yield f"jitter12_2048_strided_{tag}", dict(
            cosines=g.cosines[::64], energies=g.energies[::64], radii=g.radii, rhos=g.rhos,
            params=g.params, h_km=g.h_km, antinu=g.antinu)


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libcp3ref.so is not built (make -C oracle ref)")
    for name, c in cases():
        P = oracle.ref_propagate(c["antinu"], c["cosines"], c["energies"], c["radii"], c["rhos"],
                                 c["h_km"], c["params"], threads=4)
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            cosines=np.asarray(c["cosines"], np.float64),
                            energies=np.asarray(c["energies"], np.float64),
                            radii=np.asarray(c["radii"], np.float64), rhos=np.asarray(c["rhos"], np.float64),
                            params=np.asarray(c["params"], np.float64), h_km=np.float64(c["h_km"]),
                            antinu=np.int32(c["antinu"]), P=P)
        print(f"{name}: {P.shape[0]}x{P.shape[1]} pairs")


if __name__ == "__main__":
    main()
