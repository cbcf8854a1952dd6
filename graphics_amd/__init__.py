"""MI355X-native SPH particle loops (neighbour search, WCSPH, DFSPH, IISPH).

The compute path is libsphgpu.so: hand-written gfx950 HIP kernels behind the
C-ABI of include/sphgpu.h.  This package holds its build script, a ctypes
binding and the synthetic workloads.
"""
from . import sphgpu  # noqa: F401
from .sphgpu import Simulation, SphError, device_available  # noqa: F401

__all__ = ["sphgpu", "Simulation", "SphError", "device_available"]
