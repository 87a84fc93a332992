"""Host side of the mixed kernels' real basis (cudaprob3_amd/csrc/host_util.hpp real_basis),
on the CPU: the C++ function, compiled here with the host compiler, against its defining
properties -- W = diag(1, W') unitary, W^dagger M W real symmetric, W Hr W^dagger = M for the
vacuum part M = U diag(0, DM(1,0), DM(2,0)) U^dagger -- and against the numpy construction the
mixed-algebra emulation uses.  Cases: a CP-violating PDG matrix, the same with random row and
column phases, theta13 = 0, and theta12 = theta13 = 0 (M_10 = M_20 = 0, the branch that keeps
W' = 1 for its first column)."""
import os
import subprocess

import numpy as np
import pytest

from cudaprob3_amd.propagator import Propagator
from test_mixed_emulation import real_basis

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "real_basis_check.cpp")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("rb") / "real_basis_check")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__",
                        "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"), "-o", exe, SRC],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("host compiler cannot build the checker: " + r.stderr[-300:])
    return exe


def run(exe, U, dm):
    U18 = np.stack([U.real, U.imag], -1).reshape(18)
    inp = " ".join(f"{x:.17g}" for x in list(U18) + list(dm))
    r = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True)
    v = np.array([float(x) for x in r.stdout.split()])
    Hr = np.array([[v[0], v[3], v[4]], [v[3], v[1], v[5]], [v[4], v[5], v[2]]])
    return Hr, v[6] + 1j * v[7], v[8] + 1j * v[9]


def pdg(t12, t13, t23, dcp, dm21=7.42e-5, dm32=2.494e-3):
    p = Propagator(1, 1)
    p.setMNSMatrix(t12, t13, t23, dcp)
    p.setNeutrinoMasses(dm21, dm32)
    U18 = np.asarray(p.Mix_U, dtype=np.float64)
    return (U18[0::2] + 1j * U18[1::2]).reshape(3, 3), np.asarray(p.dm, dtype=np.float64)


CASES = {
    "pdg_dcp": pdg(0.5836, 0.1496, 0.7222, 3.86),
    "pdg_io": pdg(0.5836, 0.1496, 0.7222, 1.2, dm32=-2.494e-3),
    "theta13_zero": pdg(0.5836, 0.0, 0.7222, 3.86),
    "theta12_theta13_zero": pdg(0.0, 0.0, 0.7222, 3.86),
}


@pytest.mark.parametrize("rephase", [False, True])
@pytest.mark.parametrize("case", sorted(CASES))
def test_real_basis_host(checker, case, rephase):
    U, dm = CASES[case]
    if rephase:
        rng = np.random.default_rng(5)
        U = np.diag(np.exp(1j * rng.uniform(0, 6.3, 3))) @ U @ np.diag(np.exp(1j * rng.uniform(0, 6.3, 3)))
    Hr, w11, w12 = run(checker, U, dm)
    M = sum(dm[3 * n] * np.outer(U[:, n], np.conj(U[:, n])) for n in (1, 2))
    scale = np.abs(M).max()
    # the numpy construction of the emulation: same W row mu and the same Hr
    Hn, Wn = real_basis(U, dm)
    assert np.abs(Hr - Hn.real).max() < 1e-14 * scale
    assert abs(w11 - Wn[1, 1]) < 1e-14
    if abs(Hr[1, 2]) > 1e-10 * scale:
        # W's second column is fixed up to sign by the realness of (W^dagger M W)_12; with that
        # entry zero (theta13 = 0) any phase is a valid choice and rounding picks one
        assert abs(w12 - Wn[1, 2]) < 1e-14
    # defining properties that rows e and mu of W fix: Hr has M's spectrum (W unitary), row mu
    # is a unit vector orthogonal to row e = e_0, and (W Hr W^dagger)_ab = M_ab for a, b in (e, mu)
    assert np.abs(np.linalg.eigvalsh(Hr) - np.linalg.eigvalsh(M)).max() < 1e-14 * scale
    assert np.isclose(abs(w11) ** 2 + abs(w12) ** 2, 1.0, atol=1e-14)
    R = np.array([[1, 0, 0], [0, w11, w12]])
    assert np.abs(R @ Hr @ np.conj(R.T) - M[:2, :2]).max() < 1e-14 * scale
