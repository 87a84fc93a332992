"""MI355X-native propagation of atmospheric neutrinos through a layered Earth.

The hot path -- per (cosZ, E) pair: layer path set-up, the Barger step per crossed layer,
the running product of 3x3 transition matrices and the write of the nine P(alpha->beta)
-- runs in hand-written gfx950 HIP kernels (libcp3hip.so, C ABI in include/cp3hip.h).
``HipPropagatorSingle`` mirrors the reference's ``Propagator<FLOAT_T>`` API.
"""
from .propagator import (  # noqa: F401
    Antineutrino, HipPropagatorSingle, Neutrino, Propagator,
    e_e, e_m, e_t, m_e, m_m, m_t, t_e, t_m, t_t,
)
from . import workloads  # noqa: F401
