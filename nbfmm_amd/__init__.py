"""nbfmm_amd: MI355X-native nbfmm force evaluation and Verlet time step.

Python front end of libnbfmm_amd.so (C ABI in include/nbfmm_amd.h).  The
`Solver` class mirrors nbfmm::Solver (include/nbfmm/core/solver.hpp:22-209):
same constructor arguments, `solve`, the eight stage methods and access to
the solver's buffers in the reference layouts (what NBFMM_CHECK exposes).
Arrays are torch tensors on the GPU (float32 positions [N, 2], weights [N]).

There is no CPU fallback: if the HIP library is missing or no GPU is
present, construction raises.
"""
from __future__ import annotations

from ._lib import NbfmmError, lib, lib_path  # noqa: F401
from .solver import (BUFFERS, Solver, update, FLAG_OUT_OF_BOX, FLAG_P2P_WAIT)  # noqa: F401
from .batch import Batch  # noqa: F401

__all__ = ["Solver", "Batch", "update", "NbfmmError", "lib", "lib_path", "BUFFERS", "FLAG_OUT_OF_BOX",
           "FLAG_P2P_WAIT"]
