"""Reading side of the shuffle: a partition's Arrow IPC stream back into device
batches for the next stage.

Mirrors, for uncompressed streams, ``SchemaAwareArrowCompressedStreamReader``
(core/src/main/scala/com/intel/oap/vectorized/ArrowColumnarBatchSerializer.scala)
with ``ShuffleDecompressionJniWrapper.make`` / ``decompress``
(cpp/src/jni/jni_wrapper.cc:1236-1322): the schema is given once, then every
record batch of the stream comes back as buffers in the JNI order.  Buffers of
uncompressed batches point into the device copy of the stream itself (the read
moves no body bytes); buffers of LZ4_FRAME compressed batches (the Spark "lz4"
codec) are decoded on the device into an arena the batches then point into.
"""
from __future__ import annotations

import ctypes
from typing import List, Sequence, Tuple

import numpy as np

from . import _lib as L
from . import types as T
from .batch import DeviceBatch


class StreamBatch(DeviceBatch):
    """A DeviceBatch whose buffers are views into a device copy of an IPC stream
    (``owner`` keeps that copy alive)."""

    def __init__(self, num_rows: int, type_ids: Sequence[int], addrs, sizes, owner):
        self.num_rows = int(num_rows)
        self.type_ids = list(type_ids)
        self.tensors = [owner]
        self.addrs = (ctypes.c_int64 * max(1, len(addrs)))(*[int(a) for a in addrs])
        self.sizes = (ctypes.c_int64 * max(1, len(sizes)))(*[int(s) for s in sizes])
        self.num_bufs = len(addrs)


def parse_stream(stream: np.ndarray, device_base: int, types: Sequence[int], arena_base: int = 0
                 ) -> Tuple[np.ndarray, np.ndarray, np.ndarray, int]:
    """Host parse of one IPC stream: (rows per batch, buffer addresses, buffer
    sizes, decode arena bytes), addresses and sizes shaped (batches, buffers per
    batch).  Uncompressed buffers are offset by ``device_base``, compressed ones
    by ``arena_base`` (0 gives offsets; empty buffers give 0)."""
    stream = np.ascontiguousarray(stream, np.uint8)
    nbufs = sum(3 if T.is_binary(t) else 2 for t in types)
    tarr = L.i32_array(types)
    nb = ctypes.c_int32()
    arena = ctypes.c_int64()
    L.check(L.lib().gz_ipc_read_batches(stream.ctypes.data_as(ctypes.c_void_p), len(stream), device_base,
                                        arena_base, len(types), tarr, 0, None, None, None, ctypes.byref(nb),
                                        ctypes.byref(arena)))
    n = nb.value
    rows = np.zeros(max(1, n), np.int64)
    addrs = np.zeros((max(1, n), max(1, nbufs)), np.int64)
    sizes = np.zeros((max(1, n), max(1, nbufs)), np.int64)
    p64 = ctypes.POINTER(ctypes.c_int64)
    L.check(L.lib().gz_ipc_read_batches(stream.ctypes.data_as(ctypes.c_void_p), len(stream), device_base,
                                        arena_base, len(types), tarr, n, rows.ctypes.data_as(p64),
                                        addrs.ctypes.data_as(p64), sizes.ctypes.data_as(p64), ctypes.byref(nb),
                                        ctypes.byref(arena)))
    return rows[:n], addrs[:n, :nbufs], sizes[:n, :nbufs], arena.value


class ShuffleReader:
    """``ShuffleDecompressionJniWrapper`` for one schema: ``read`` turns one
    partition's stream into device batches (``ctx``: the context whose stream
    runs the LZ4 decoding; needed for compressed streams only)."""

    def __init__(self, types: Sequence[int], ctx=None):
        self.types = list(types)
        self.ctx = ctx

    def _batches(self, stream: np.ndarray, device, base: int) -> List[StreamBatch]:
        import torch

        rows, addrs, sizes, arena_bytes = parse_stream(stream, base, self.types)
        owner = device
        if arena_bytes:
            if self.ctx is None:
                raise L.Invalid(4, "a compressed stream needs a context to decode on")
            arena = torch.empty(arena_bytes, dtype=torch.uint8, device=device.device)
            rows, addrs, sizes, _ = parse_stream(stream, base, self.types, int(arena.data_ptr()))
            L.check(L.lib().gz_ipc_decompress(self.ctx.handle, stream.ctypes.data_as(ctypes.c_void_p), len(stream),
                                              base, len(self.types), L.i32_array(self.types),
                                              ctypes.c_void_p(arena.data_ptr()), arena_bytes))
            owner = (device, arena)
        return [StreamBatch(rows[b], self.types, addrs[b], sizes[b], owner) for b in range(len(rows))]

    def read(self, stream: np.ndarray, device=None) -> List[StreamBatch]:
        """``stream``: the host bytes of one partition's stream.  ``device``: the
        same bytes already in HBM (a torch uint8 tensor), or None to upload them."""
        import torch

        stream = np.ascontiguousarray(stream, np.uint8)
        if device is None:
            device = torch.from_numpy(stream.copy()).to("cuda")
        if device.numel() < len(stream):
            raise L.Invalid(4, "device copy shorter than the stream")
        return self._batches(stream, device, int(device.data_ptr()))

    def read_data_file(self, data: np.ndarray, lengths: Sequence[int], device=None
                       ) -> List[List[StreamBatch]]:
        """Every partition of a Splitter data file (partition streams back to back,
        ``lengths`` bytes each); the whole file is uploaded once when ``device``
        is None.  Empty partitions give no batches."""
        import torch

        data = np.ascontiguousarray(data, np.uint8)
        if device is None:
            device = torch.from_numpy(data.copy()).to("cuda")
        out, pos = [], 0
        for n in lengths:
            n = int(n)
            if n == 0:
                out.append([])
                continue
            out.append(self._batches(data[pos:pos + n], device, int(device.data_ptr()) + pos))
            pos += n
        return out
