"""Columnar batches in the JNI buffer convention of the reference.

A batch is a list of columns; each column holds the Arrow buffers the reference's
JNI layer hands to native code (``jni_common.h:309-340``): an optional validity
bitmap, an int32 offsets buffer for utf8/binary, and the values (a bitmap for
booleans).  Buffers start at row 0 (no Arrow slice offset).

``HostBatch`` keeps numpy buffers; ``DeviceBatch`` keeps torch tensors in HBM and
yields the (addresses, sizes) arrays the C ABI takes.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import types as T


@dataclass
class HostColumn:
    type_id: int
    values: Optional[np.ndarray]          # fixed: typed array; bool: uint8 bitmap; binary: uint8 data
    validity: Optional[np.ndarray] = None  # uint8 bitmap or None
    offsets: Optional[np.ndarray] = None   # int32 (n + 1) for utf8/binary


@dataclass
class HostBatch:
    num_rows: int
    columns: List[HostColumn] = field(default_factory=list)

    @property
    def types(self) -> List[int]:
        return [c.type_id for c in self.columns]

    def buffers(self) -> List[Optional[np.ndarray]]:
        """Flat buffer list in the JNI order (validity, [offsets], values)."""
        out: List[Optional[np.ndarray]] = []
        for c in self.columns:
            out.append(c.validity)
            if T.is_binary(c.type_id):
                out.append(c.offsets)
            out.append(c.values)
        return out


def bitmap_from_bools(valid: np.ndarray) -> np.ndarray:
    """LSB-first Arrow bitmap, padded with 1 bits to a multiple of 64 bits."""
    n = len(valid)
    nwords = (n + 63) // 64
    bits = np.ones(nwords * 64, dtype=np.uint8)
    bits[:n] = valid.astype(np.uint8)
    return np.packbits(bits, bitorder="little")


def bools_from_bitmap(bm: Optional[np.ndarray], n: int) -> np.ndarray:
    if bm is None:
        return np.ones(n, dtype=bool)
    return np.unpackbits(np.asarray(bm, dtype=np.uint8), bitorder="little")[:n].astype(bool)


# ---------------------------------------------------------------- pyarrow bridge

def from_arrow(rb) -> HostBatch:
    """pyarrow.RecordBatch -> HostBatch (copies, normalizing slice offsets)."""
    import pyarrow as pa

    cols = []
    n = rb.num_rows
    for arr in rb.columns:
        tid = T.from_arrow_type(arr.type)
        if arr.offset != 0:
            arr = pa.concat_arrays([arr])  # re-materialize at offset 0
        bufs = arr.buffers()
        validity = None
        if tid == T.NA:
            cols.append(HostColumn(tid, None, None, None))
            continue
        if arr.null_count > 0 and bufs[0] is not None:
            validity = bitmap_from_bools(
                np.asarray(arr.is_valid().to_numpy(zero_copy_only=False), dtype=bool))
        if T.is_binary(tid):
            offsets = np.frombuffer(bufs[1], dtype=np.int32, count=n + 1).copy()
            start, end = int(offsets[0]), int(offsets[-1])
            data = np.frombuffer(bufs[2], dtype=np.uint8)[start:end].copy() if bufs[2] is not None \
                else np.zeros(0, np.uint8)
            offsets -= start
            cols.append(HostColumn(tid, data, validity, offsets))
        elif tid == T.BOOL:
            vals = np.asarray(arr.fill_null(False).to_numpy(zero_copy_only=False), dtype=bool)
            cols.append(HostColumn(tid, bitmap_from_bools(vals), validity))
        else:
            width = T.width(tid)
            raw = np.frombuffer(bufs[1], dtype=np.uint8, count=n * width).copy()
            cols.append(HostColumn(tid, raw.view(T.numpy_dtype(tid)), validity))
    return HostBatch(n, cols)


def column_to_arrow(col: HostColumn, n: int, arrow_type=None):
    import pyarrow as pa

    typ = arrow_type if arrow_type is not None else T.to_arrow_type(col.type_id)
    if col.type_id == T.NA:
        return pa.nulls(n)
    validity = None
    null_count = 0
    if col.validity is not None:
        valid = bools_from_bitmap(col.validity, n)
        null_count = int(n - valid.sum())
        validity = pa.py_buffer(bitmap_from_bools(valid)) if null_count else None
    if T.is_binary(col.type_id):
        return pa.Array.from_buffers(typ, n, [validity, pa.py_buffer(np.ascontiguousarray(col.offsets)),
                                             pa.py_buffer(np.ascontiguousarray(col.values))],
                                     null_count=null_count)
    return pa.Array.from_buffers(typ, n, [validity, pa.py_buffer(np.ascontiguousarray(col.values))],
                                 null_count=null_count)


def to_arrow(batch: HostBatch, schema=None):
    import pyarrow as pa

    arrays = [column_to_arrow(c, batch.num_rows, schema.field(i).type if schema is not None else None)
              for i, c in enumerate(batch.columns)]
    if schema is None:
        return pa.RecordBatch.from_arrays(arrays, names=[f"c{i}" for i in range(len(arrays))])
    return pa.RecordBatch.from_arrays(arrays, schema=schema)


# ---------------------------------------------------------------- device side

class DeviceBatch:
    """A batch whose buffers live in HBM as torch uint8 tensors."""

    def __init__(self, num_rows: int, type_ids: Sequence[int], tensors: List[Optional["object"]]):
        self.num_rows = int(num_rows)
        self.type_ids = list(type_ids)
        self.tensors = tensors  # flat JNI-order list (None for absent buffers)
        addrs = [0 if t is None or t.numel() == 0 else int(t.data_ptr()) for t in tensors]
        sizes = [0 if t is None else int(t.numel() * t.element_size()) for t in tensors]
        self.addrs = (ctypes.c_int64 * max(1, len(addrs)))(*addrs)
        self.sizes = (ctypes.c_int64 * max(1, len(sizes)))(*sizes)
        self.num_bufs = len(addrs)

    @classmethod
    def from_host(cls, batch: HostBatch, device="cuda") -> "DeviceBatch":
        import torch

        tensors = []
        for buf in batch.buffers():
            if buf is None:
                tensors.append(None)
            else:
                a = np.ascontiguousarray(buf).view(np.uint8).reshape(-1)
                tensors.append(torch.from_numpy(a.copy()).to(device))
        return cls(batch.num_rows, batch.types, tensors)
