"""gpsm_amd — GpSM's filter-and-join subgraph matcher, built for MI355X (gfx950).

The matcher itself is native: HIP kernels and a C ABI in libgpsm.so
(include/gpsm.h).  This package loads that library with ctypes and mirrors
the reference's host interface for Python callers (tests, benchmarks).
"""
from ._lib import GpsmError, lib  # noqa: F401
from .graphs import DataGraph, Query  # noqa: F401
