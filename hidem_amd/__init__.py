"""MI355X (gfx950) implementation of HiDEM's per-timestep hot path.

The compute path is libhidem_amd.so (HIP kernels behind the C ABI of
include/hidem/hdm_abi.h).  This Python package only carries the ctypes
bindings (_abi) and the TOML parameter derivation (params) used by tests and
bench.py; importing it never falls back to a CPU path.
"""
from . import _abi  # noqa: F401

__all__ = ["_abi"]
