"""Device context: one GPU, one HIP stream, one HBM budget."""
from __future__ import annotations

import ctypes
from typing import Optional

from . import _lib as L


class Context:
    """Owns a gz_context (include/gazelle_amd.h).  ``stream`` may be a torch
    stream handle (``torch.cuda.current_stream().cuda_stream``) so that library
    launches are ordered with torch work and visible to torch events."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, stream: Optional[int] = None):
        self._h = ctypes.c_void_p()
        # Inputs are read by work queued on the context's stream after a call
        # returns (stream-ordered borrowing): the device batches passed in are
        # held here until the stream is synchronized, so their memory cannot be
        # reused under a running kernel.
        self._held = []
        L.check(L.lib().gz_context_create(device, hbm_budget, ctypes.byref(self._h)))
        if stream is not None:
            self.set_stream(stream)

    @property
    def handle(self):
        return self._h

    @property
    def alive(self) -> bool:
        """False once destroyed: the handles made on it were freed with it
        and must not be released again."""
        return bool(self._h)

    def set_stream(self, stream: Optional[int]):
        L.check(L.lib().gz_context_set_stream(self._h, ctypes.c_void_p(stream or 0)))

    def hold(self, *objs):
        """Keep input batches alive until the next synchronize()."""
        self._held.extend(objs)

    def synchronize(self):
        L.check(L.lib().gz_context_synchronize(self._h))
        self._held.clear()

    def bytes_in_use(self) -> int:
        return L.lib().gz_context_bytes_in_use(self._h)

    def peak_bytes(self) -> int:
        return L.lib().gz_context_peak_bytes(self._h)

    def set_profiling(self, on: bool):
        L.check(L.lib().gz_context_set_profiling(self._h, int(on)))

    def kernel_time(self, name: str):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        L.check(L.lib().gz_context_kernel_time(self._h, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def reset_kernel_times(self):
        L.check(L.lib().gz_context_reset_kernel_times(self._h))

    def close(self):
        if self._h:
            L.lib().gz_context_destroy(self._h)
            self._h = ctypes.c_void_p()
        self._held = []

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
