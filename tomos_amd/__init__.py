"""tomos_amd -- MI355X-native LSQR hot path of TOMOS.

The product is libtomos_hip.so (HIP kernels for gfx950 behind the C ABI in
include/tomos_hip.h) and libtomos.so (the C++ mirror of TOMOS's Dense /
Workspace / dispatch-registry / lsqr_solve API, include/tomos/*.hpp).  This
Python package binds the C ABI with ctypes for tests, benchmarks and the
torch.autograd adjoint.
"""
from . import _capi  # noqa: F401
from ._capi import DeviceBuffer, TomosError, lib  # noqa: F401

__all__ = ["DeviceBuffer", "TomosError", "lib"]
