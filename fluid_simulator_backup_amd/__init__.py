"""MI355X-native timestep hot path of a two-phase level-set fluid solver.

The product is the C-ABI library ``libfs_hip.so`` (include/fs_abi.h) built
from ``csrc/``; this package locates it and wraps it with ctypes.  There is
no CPU fallback: ``load_library()`` raises when the HIP build is missing.
"""
import os

from . import _abi
from ._abi import FsError, Solver, default_params, unit_cube_grid  # noqa: F401

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libfs_hip.so")

_lib = None


def load_library():
    """Load the HIP library (raises FileNotFoundError if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        _lib = _abi.load(LIB_PATH, "fs_")
    return _lib


def create_solver(params, grid, device=0):
    """A solver handle on the GPU (EULERIAN_FLUID_SOLVER_3D equivalent)."""
    return Solver(load_library(), "fs_", params, grid, device)
