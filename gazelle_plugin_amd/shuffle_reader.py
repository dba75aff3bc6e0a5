"""Reading side of the shuffle: a partition's Arrow IPC stream back into device
batches for the next stage.

Mirrors, for uncompressed streams, ``SchemaAwareArrowCompressedStreamReader``
(core/src/main/scala/com/intel/oap/vectorized/ArrowColumnarBatchSerializer.scala)
with ``ShuffleDecompressionJniWrapper.make`` / ``decompress``
(cpp/src/jni/jni_wrapper.cc:1236-1322): the schema is given once, then every
record batch of the stream comes back as buffers in the JNI order.  Decompression
of an UNCOMPRESSED stream is the identity, so the batches here point into the
device copy of the stream itself and the read moves no body bytes.
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


def parse_stream(stream: np.ndarray, device_base: int, types: Sequence[int]
                 ) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Host parse of one IPC stream: (rows per batch, buffer addresses, buffer
    sizes), the last two shaped (batches, buffers per batch), addresses offset
    by ``device_base`` (0 gives offsets into the stream, 0 for empty buffers)."""
    stream = np.ascontiguousarray(stream, np.uint8)
    nbufs = sum((3 if T.is_binary(t) else 2) for t in types)
    tarr = L.i32_array(types)
    nb = ctypes.c_int32()
    L.check(L.lib().gz_ipc_read_batches(stream.ctypes.data_as(ctypes.c_void_p), len(stream), device_base,
                                        len(types), tarr, 0, None, None, None, ctypes.byref(nb)))
    n = nb.value
    rows = np.zeros(max(1, n), np.int64)
    addrs = np.zeros((max(1, n), max(1, nbufs)), np.int64)
    sizes = np.zeros((max(1, n), max(1, nbufs)), np.int64)
    p64 = ctypes.POINTER(ctypes.c_int64)
    L.check(L.lib().gz_ipc_read_batches(stream.ctypes.data_as(ctypes.c_void_p), len(stream), device_base,
                                        len(types), tarr, n, rows.ctypes.data_as(p64), addrs.ctypes.data_as(p64),
                                        sizes.ctypes.data_as(p64), ctypes.byref(nb)))
    return rows[:n], addrs[:n, :nbufs], sizes[:n, :nbufs]


class ShuffleReader:
    """``ShuffleDecompressionJniWrapper`` for one schema: ``read`` turns one
    partition's stream into device batches."""

    def __init__(self, types: Sequence[int]):
        self.types = list(types)

    def read(self, stream: np.ndarray, device=None) -> List[StreamBatch]:
        """``stream``: the host bytes of one partition's stream.  ``device``: the
        same bytes already in HBM (a torch uint8 tensor), or None to upload them."""
        import torch

        stream = np.ascontiguousarray(stream, np.uint8)
        if device is None:
            device = torch.from_numpy(stream.copy()).to("cuda")
        if device.numel() < len(stream):
            raise L.Invalid(4, "device copy shorter than the stream")
        rows, addrs, sizes = parse_stream(stream, int(device.data_ptr()), self.types)
        return [StreamBatch(rows[b], self.types, addrs[b], sizes[b], device) for b in range(len(rows))]

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
            rows, addrs, sizes = parse_stream(data[pos:pos + n], int(device.data_ptr()) + pos, self.types)
            out.append([StreamBatch(rows[b], self.types, addrs[b], sizes[b], device) for b in range(len(rows))])
            pos += n
        return out
