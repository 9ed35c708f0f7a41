"""MI355X-native time-dependent wavepacket propagator.

The compute path is libqd_hip.so (hand-written HIP kernels for gfx950 behind
the C ABI of include/qd.h).  This package is the host side: problem setup
(``problem``), the ctypes binding (``_abi``) and the mirror of the
reference's entry points (``propagator``).
"""
from . import problem  # noqa: F401
from .propagator import Propagator, TimeEvolutionMex, TimeEvolutionMexCUDA, run  # noqa: F401
