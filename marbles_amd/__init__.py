"""marbles_amd: MI355X-native lattice-Boltzmann engine for the MARBLES hot path.

The compute lives in libmarbles_amd.so (hand-written HIP for gfx950 behind a
C ABI, include/marbles_amd.h).  This package is the Python host surface:
`Level` mirrors the per-level calls of the reference solver and `configs`
describes the standard workloads with the reference's input-deck keys.
There is no CPU fallback: using a Level without the built library raises.
"""
from ._lib import MrbError, load  # noqa: F401
from .level import Level  # noqa: F401
from . import configs  # noqa: F401

__all__ = ["Level", "MrbError", "load", "configs"]
from .amr import Hierarchy  # noqa: F401,E402
