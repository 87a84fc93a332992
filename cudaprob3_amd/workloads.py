"""Synthetic inputs of the propagation workloads (seeded, reproducible).

Earth models are the public PREM profiles that the reference ships as two-column
density files (example/models/PREM_{4,12,59}layer.dat: radius in km ascending from
r = 0, density in g/cm^3; the density of row j applies between radius j-1 and j).

Workloads (names used by bench.py and the tests):

* ``prem4_256``   -- 256 cosZ = -1 + i/256 x 256 energies 10^(-1+3j/255) GeV,
  PREM 4-layer, production height 22 km, neutrino mode.
* ``jitter12_2048`` -- 2048 cosZ in [-1, 1] x 2048 energies log-spaced 0.05-1000 GeV
  on a 12-shell PREM-like Earth with interior radii jittered by +-2 % (seed 2),
  densities non-increasing from 13.0 to 2.6 g/cm^3; neutrino/antineutrino x normal
  /inverted ordering.
* ``scan1024``    -- 1024 parameter sets with uneven grids and their own random
  monotone Earth models of 4-50 shells (seed 3).
* ``fine100_8192`` -- 8192 x 8192 grid on 100 equal-thickness shells sampled from the
  59-row PREM profile at shell midpoints (seed 5 for the Markov chain).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

R_EARTH_KM = 6371.0

# example/models/PREM_4layer.dat
PREM4 = np.array([[0.0, 13.0], [1220.0, 13.0], [3480.0, 11.3], [5701.0, 5.0], [6371.0, 3.3]])

# example/models/PREM_12layer.dat
PREM12 = np.array([
    [0.00, 13.0], [1221.50, 13.0], [1974.33, 11.948163], [2727.17, 11.341572],
    [3480.00, 10.462371], [4220.33, 5.381883], [4960.67, 5.007959], [5701.00, 4.603493],
    [5971.00, 3.9], [6346.60, 3.4], [6356.00, 2.900000], [6368.00, 2.600000], [6371.00, 1.020000]])

# example/models/PREM_59layer.dat
PREM59 = np.array([
    [0.0, 13.088500], [100.0, 13.087956], [200.0, 13.083601], [300.0, 13.074891],
    [400.0, 13.061827], [500.0, 13.044407], [600.0, 13.022633], [700.0, 12.996504],
    [800.0, 12.966020], [900.0, 12.931181], [1000.0, 12.891987], [1100.0, 12.848439],
    [1200.0, 12.800535], [1221.5, 12.769308], [1400.0, 12.119166], [1500.0, 12.040013],
    [1600.0, 11.978818], [1700.0, 11.913841], [1800.0, 11.844952], [1900.0, 11.772024],
    [2000.0, 11.694929], [2100.0, 11.613537], [2200.0, 11.527721], [2300.0, 11.437353],
    [2400.0, 11.342303], [2500.0, 11.242445], [2600.0, 11.137649], [2700.0, 11.027788],
    [2800.0, 10.912733], [2900.0, 10.792356], [3000.0, 10.666529], [3100.0, 10.535122],
    [3200.0, 10.398009], [3300.0, 10.255061], [3400.0, 10.106150], [3480.0, 9.966925],
    [3600.0, 5.536414], [3630.0, 5.498957], [3800.0, 5.449133], [4000.0, 5.357086],
    [4200.0, 5.257316], [4400.0, 5.156720], [4600.0, 5.054725], [4800.0, 4.950760],
    [5000.0, 4.844253], [5200.0, 4.734632], [5400.0, 4.621326], [5600.0, 4.503762],
    [5701.0, 4.412141], [5771.0, 3.983971], [5871.0, 3.912802], [5971.0, 3.786765],
    [6061.0, 3.516391], [6151.0, 3.462647], [6221.0, 3.363294], [6291.0, 3.370902],
    [6346.6, 3.377727], [6356.0, 2.900000], [6368.0, 2.600000], [6371.0, 1.020000]])

# Oscillation parameters of the 256x256 and 2048x2048 workloads:
# theta12, theta13, theta23, deltaCP (rad), dm21^2, dm32^2 (eV^2)
NUFIT = (0.5836, 0.1496, 0.7222, 3.86, 7.42e-5, 2.494e-3)
PRODUCTION_HEIGHT_KM = 22.0


@dataclass
class Grid:
    cosines: np.ndarray
    energies: np.ndarray
    radii: np.ndarray            # density-file order (ascending from r = 0)
    rhos: np.ndarray
    params: tuple = NUFIT
    antinu: int = 0
    h_km: float = PRODUCTION_HEIGHT_KM
    name: str = ""

    @property
    def n_pairs(self) -> int:
        return len(self.cosines) * len(self.energies)


def prem4_256() -> Grid:
    cos = -1.0 + np.arange(256) / 256.0
    E = 10.0 ** (-1.0 + 3.0 * np.arange(256) / 255.0)
    return Grid(cos, E, PREM4[:, 0].copy(), PREM4[:, 1].copy(), name="prem4_256")


def jitter12_model(seed: int = 2) -> tuple[np.ndarray, np.ndarray]:
    """12 shells around PREM_12 radii, interior radii x (1 + U(-0.02, 0.02))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = PREM12[:, 0].copy()
    interior = r[1:-1] * (1.0 + rng.uniform(-0.02, 0.02, size=len(r) - 2))
    interior = np.sort(np.clip(interior, 1.0, R_EARTH_KM - 1.0))
    # keep every shell at least 0.5 km thick (the outer radii sit within 2 % of R)
    upper = R_EARTH_KM
    for i in range(len(interior) - 1, -1, -1):
        interior[i] = min(interior[i], upper - 0.5)
        upper = interior[i]
    radii = np.concatenate([[0.0], interior, [R_EARTH_KM]])
    rho = PREM12[:, 1].copy()
    rho[rho < 2.6] = 2.6
    rho = np.minimum.accumulate(rho)          # non-increasing outward, 13.0 -> 2.6
    return radii, rho


def jitter12_2048(n: int = 2048, antinu: int = 0, inverted: bool = False) -> Grid:
    cos = np.linspace(-1.0, 1.0, n)
    E = np.geomspace(0.05, 1000.0, n)
    radii, rho = jitter12_model()
    p = list(NUFIT)
    if inverted:
        p[5] = -p[5]
    return Grid(cos, E, radii, rho, params=tuple(p), antinu=antinu, name=f"jitter12_{n}")


def jitter12_variants(n: int = 2048) -> list[Grid]:
    """nu/nubar x normal/inverted ordering on the jittered 12-shell Earth."""
    return [jitter12_2048(n, antinu=a, inverted=inv) for inv in (False, True) for a in (0, 1)]


def fine100_model() -> tuple[np.ndarray, np.ndarray]:
    """100 equal-thickness shells, densities of PREM_59 (step profile) at shell midpoints."""
    edges = np.linspace(0.0, R_EARTH_KM, 101)
    mids = 0.5 * (edges[:-1] + edges[1:])
    # density of a point at radius r: row j with r_{j-1} < r <= r_j
    idx = np.searchsorted(PREM59[:, 0], mids, side="left")
    rho_shell = PREM59[idx, 1]
    radii = edges
    rhos = np.concatenate([[rho_shell[0]], rho_shell])
    return radii, rhos


def fine100_grid(n: int = 8192) -> Grid:
    radii, rhos = fine100_model()
    return Grid(np.linspace(-1.0, 1.0, n), np.geomspace(0.1, 1000.0, n), radii, rhos, name=f"fine100_{n}")


PARAM_LO = np.array([0.55, 0.14, 0.65, 0.0, 6.8e-5, 2.3e-3])
PARAM_HI = np.array([0.62, 0.16, 0.92, 2 * np.pi, 8.0e-5, 2.7e-3])


@dataclass
class ScanSet:
    params: tuple
    cosines: np.ndarray
    energies: np.ndarray
    radii: np.ndarray
    rhos: np.ndarray
    antinu: int = 0
    h_km: float = PRODUCTION_HEIGHT_KM


def random_params(rng: np.random.Generator) -> tuple:
    p = rng.uniform(PARAM_LO, PARAM_HI)
    sign = 1.0 if rng.random() < 0.5 else -1.0
    p[5] *= sign
    return tuple(float(x) for x in p)


def random_earth(rng: np.random.Generator, n_shells: int) -> tuple[np.ndarray, np.ndarray]:
    inner = np.sort(rng.uniform(0.0, R_EARTH_KM, size=n_shells - 1))
    radii = np.concatenate([[0.0], inner, [R_EARTH_KM]])
    dens = np.sort(rng.uniform(2.5, 13.0, size=n_shells))[::-1]   # centre -> surface, descending
    rhos = np.concatenate([[dens[0]], dens])
    return radii, rhos


def scan_sets(n_sets: int = 1024, seed: int = 3, nc_range=(64, 512), ne_range=(64, 1024),
              shells=(4, 50)) -> list[ScanSet]:
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n_sets):
        p = random_params(rng)
        n_c = int(rng.integers(nc_range[0], nc_range[1] + 1))
        n_e = int(rng.integers(ne_range[0], ne_range[1] + 1))
        k = int(rng.integers(shells[0], shells[1] + 1))
        radii, rhos = random_earth(rng, k)
        out.append(ScanSet(p, np.linspace(-1.0, 1.0, n_c), np.geomspace(0.1, 100.0, n_e), radii, rhos))
    return out
