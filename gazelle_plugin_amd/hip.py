"""Minimal HIP runtime access for host-side readback of device results."""
from __future__ import annotations

import ctypes

import numpy as np

_hip = None
H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        from ._lib import _bind_process_hip_runtime

        _bind_process_hip_runtime()
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise ImportError("libamdhip64.so not found")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
        _hip.hipDeviceSynchronize.restype = ctypes.c_int
    return _hip


def to_host(dev_ptr, nbytes: int, dtype=np.uint8) -> np.ndarray:
    """Copies nbytes from a device address into a new numpy array."""
    out = np.empty(max(0, int(nbytes)), dtype=np.uint8)
    if nbytes > 0:
        if not dev_ptr:
            raise ValueError("null device pointer")
        rc = hip().hipMemcpy(out.ctypes.data, ctypes.c_void_p(dev_ptr), int(nbytes), D2H)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed: {rc}")
    return out.view(dtype)
