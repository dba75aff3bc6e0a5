"""Python mirror of the reference shuffle Splitter over the C ABI.

Mirrors ``sparkcolumnarplugin::shuffle::Splitter`` (cpp/src/shuffle/splitter.h:38-295):
``Splitter.Make(short_name, schema, num_partitions, key_cols, options)``, ``Split``,
``Stop``, ``PartitionLengths``.  Errors raise ``GazelleError`` subclasses carrying
the arrow::StatusCode value, where the reference returns an ``arrow::Status``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib as L
from . import types as T
from .batch import DeviceBatch
from .context import Context
from .hip import to_host


@dataclass
class SplitOptions:
    """``SplitOptions`` (cpp/src/shuffle/type.h:39-59), the fields the device path uses."""
    buffer_size: int = 4096
    prefer_spill: bool = True
    offheap_per_task: int = 0
    tile_rows: int = 0
    hash_seed: int = 42
    append_row_ids: bool = False  # last output column: int32 row index within the split call
    batch_compress_threshold: int = 256  # LZ4: record batches of at most this many rows stay uncompressed

    def to_c(self) -> L.GzSplitOptions:
        o = L.GzSplitOptions()
        o.buffer_size = self.buffer_size
        o.prefer_spill = int(self.prefer_spill)
        o.offheap_per_task = self.offheap_per_task
        o.tile_rows = self.tile_rows
        o.hash_seed = self.hash_seed
        o.append_row_ids = int(self.append_row_ids)
        o.batch_compress_threshold = self.batch_compress_threshold
        return o


class Splitter:
    def __init__(self, ctx: Context, handle, types: Sequence[int], P: int, mode: str):
        self.ctx = ctx
        self._h = handle
        self.input_types = list(types)
        self.num_partitions = P
        self.mode = mode
        self.output_types = list(types[1:]) if mode == "range" else list(types)

    @classmethod
    def Make(cls, ctx: Context, short_name: str, types: Sequence[int], num_partitions: int,
             key_cols: Sequence[int] = (), options: Optional[SplitOptions] = None) -> "Splitter":
        opts = (options or SplitOptions()).to_c()
        h = ctypes.c_void_p()
        L.check(L.lib().gz_split_make(ctx.handle, short_name.encode(), len(types), L.i32_array(types),
                                      num_partitions, L.i32_array(key_cols), len(key_cols),
                                      ctypes.byref(opts), ctypes.byref(h)))
        P = 1 if short_name == "single" else num_partitions
        sp = cls(ctx, h, types, P, short_name)
        if options is not None and options.append_row_ids:
            sp.output_types.append(T.INT32)
        return sp

    # -- splitting
    def Split(self, batch: DeviceBatch):
        """One batch, one launch sequence (the plain per-batch path)."""
        self.ctx.hold(batch)
        L.check(L.lib().gz_split_batch(self._h, batch.num_rows, batch.addrs, batch.sizes, batch.num_bufs))

    def SplitBatches(self, batches: Sequence[DeviceBatch], arr=None):
        """Batched form: all batches in one launch sequence.  ``arr`` may be a
        prebuilt ``make_batch_array(batches)``."""
        self.ctx.hold(list(batches))
        if arr is None:
            arr = make_batch_array(batches)
        L.check(L.lib().gz_split_batches(self._h, len(batches), arr))

    def Stop(self):
        L.check(L.lib().gz_split_stop(self._h))

    def PartitionLengths(self) -> np.ndarray:
        """Rows per partition (the reference reports bytes of the written file;
        rows are what the device split produces)."""
        out = np.zeros(self.num_partitions, np.int64)
        L.check(L.lib().gz_split_partition_rows(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return out

    def TotalComputePidTime(self) -> int:
        return L.lib().gz_split_compute_pid_time(self._h)

    def RawPartitionLengths(self) -> np.ndarray:
        """In-memory bytes of each partition's record batches (after the data file
        is laid out: ipc_stream / gz_split_ipc_prepare)."""
        out = np.zeros(self.num_partitions, np.int64)
        L.check(L.lib().gz_split_raw_partition_lengths(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return out

    def BatchRows(self) -> int:
        """Rows of a partition buffer (CalculateSplitBatchSize) used for the cuts."""
        n = ctypes.c_int64()
        L.check(L.lib().gz_split_batch_rows(self._h, ctypes.byref(n)))
        return n.value

    def Metrics(self) -> Dict[str, int]:
        """TotalComputePidTime / WriteTime / SpillTime / CompressTime (ns) and
        TotalBytesWritten / BytesSpilled, as the JNI stop's SplitResult has them."""
        m = L.GzSplitMetrics()
        L.check(L.lib().gz_split_get_metrics(self._h, ctypes.byref(m)))
        return {f: getattr(m, f) for f, _ in L.GzSplitMetrics._fields_}

    def num_segments(self) -> int:
        return L.lib().gz_split_num_segments(self._h)

    def segment_batch_counts(self, seg: int) -> np.ndarray:
        nb = ctypes.c_int32()
        L.check(L.lib().gz_split_segment_batch_counts(self._h, seg, ctypes.byref(nb), None))
        out = np.zeros(max(1, nb.value) * self.num_partitions, np.int32)
        L.check(L.lib().gz_split_segment_batch_counts(self._h, seg, ctypes.byref(nb),
                                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return out[: nb.value * self.num_partitions].reshape(nb.value, self.num_partitions)

    # -- readback
    def segment(self, seg: int) -> Dict:
        """Host copy of one segment's partitioned layout (drains the context's
        stream first: the copies run on the null stream)."""
        if self.ctx is not None:
            self.ctx.synchronize()
        P = self.num_partitions
        rs = np.zeros(P + 1, np.int64)
        L.check(L.lib().gz_split_segment_rows(self._h, seg, rs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        N = int(rs[-1])
        cols = []
        for c, t in enumerate(self.output_types):
            vals, valid, offs, data = (ctypes.c_void_p() for _ in range(4))
            bs = np.zeros(P + 1, np.int64)
            nulls = ctypes.c_int64()
            L.check(L.lib().gz_split_segment_column(
                self._h, seg, c, ctypes.byref(vals), ctypes.byref(valid), ctypes.byref(offs),
                ctypes.byref(data), bs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(nulls)))
            nbm = (N + 63) // 64 * 8
            col = {"type": t, "null_count": nulls.value,
                   "validity": to_host(valid.value, nbm) if valid.value else None}
            if T.is_binary(t):
                col["offsets"] = to_host(offs.value, 4 * (N + P), np.int32)
                col["byte_start"] = bs
                col["data"] = to_host(data.value, int(bs[-1]))
            elif t == T.BOOL:
                col["values"] = to_host(vals.value, nbm)
            elif t != T.NA:
                col["values"] = to_host(vals.value, N * T.width(t))
            cols.append(col)
        return {"row_start": rs, "num_rows": N, "columns": cols}

    def read_partitions(self) -> Dict:
        """Concatenation over segments, per partition, in the same layout as one
        segment (and as oracle.split produces)."""
        segs = [self.segment(s) for s in range(self.num_segments())]
        return concat_segments(segs, self.output_types, self.num_partitions)

    def SetCompressType(self, codec: str):
        """``Splitter::SetCompressType`` (splitter.cc:371-385) for the data file:
        "uncompressed" or "lz4" (LZ4_FRAME, as the reference maps Spark's lz4)."""
        codes = {"uncompressed": 0, "none": 0, "lz4": 6, "lz4_frame": 6}
        if codec.lower() not in codes:
            raise L.NotImplementedStatus(10, f"compression type {codec}")
        L.check(L.lib().gz_split_set_compress_type(self._h, codes[codec.lower()]))

    def ipc_stream(self, schema_message: bytes = b"", write_schema: bool = True):
        """The data file Splitter::Stop writes (partitions' Arrow IPC streams, one
        after another, uncompressed), assembled on the device and copied to the
        host: (stream bytes as a numpy uint8 array, partition lengths in bytes).
        ``schema_message`` is the encapsulated schema message (for example
        ``pyarrow.Schema.serialize()``), written first in every non-empty
        partition when ``write_schema``."""
        import torch

        total = ctypes.c_int64()
        lengths = np.zeros(L.lib().gz_split_num_partitions(self._h), np.int64)
        buf = np.frombuffer(schema_message, np.uint8) if schema_message else np.zeros(1, np.uint8)
        L.check(L.lib().gz_split_ipc_prepare(self._h, buf.ctypes.data_as(ctypes.c_void_p), len(schema_message),
                                             int(bool(write_schema)), ctypes.byref(total),
                                             lengths.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        dev = torch.empty(max(1, total.value), dtype=torch.uint8, device="cuda")
        L.check(L.lib().gz_split_ipc_write(self._h, ctypes.c_void_p(dev.data_ptr()), total.value))
        torch.cuda.synchronize()
        return dev[: total.value].cpu().numpy(), lengths

    def write_data_file(self, path: str, schema_message: bytes = b"", write_schema: bool = True) -> np.ndarray:
        """Splitter::Stop's data file at ``path``; returns the partition lengths."""
        data, lengths = self.ipc_stream(schema_message, write_schema)
        with open(path, "wb") as f:
            f.write(data.tobytes())
        return lengths

    def close(self):
        if self._h:
            if self.ctx.alive:
                L.lib().gz_split_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_batch_array(batches: Sequence[DeviceBatch]):
    arr = (L.GzBatch * max(1, len(batches)))()
    for i, b in enumerate(batches):
        arr[i].num_rows = b.num_rows
        arr[i].buf_addrs = ctypes.cast(b.addrs, ctypes.POINTER(ctypes.c_int64))
        arr[i].buf_sizes = ctypes.cast(b.sizes, ctypes.POINTER(ctypes.c_int64))
        arr[i].num_bufs = b.num_bufs
    return arr


def _bits(bm: Optional[np.ndarray], start: int, n: int) -> np.ndarray:
    if bm is None:
        return np.ones(n, bool)
    return np.unpackbits(bm, bitorder="little")[start:start + n].astype(bool)


def concat_segments(segs: List[Dict], types: Sequence[int], P: int) -> Dict:
    if len(segs) == 1:
        return segs[0]
    counts = np.array([np.diff(s["row_start"]) for s in segs]) if segs else np.zeros((0, P), np.int64)
    tot = counts.sum(axis=0) if len(segs) else np.zeros(P, np.int64)
    row_start = np.zeros(P + 1, np.int64)
    row_start[1:] = np.cumsum(tot)
    N = int(row_start[-1])
    cols = []
    for c, t in enumerate(types):
        col = {"type": t, "null_count": sum(s["columns"][c]["null_count"] for s in segs)}
        has_valid = any(s["columns"][c]["validity"] is not None for s in segs)
        parts_valid, parts_vals, parts_bool, parts_data, parts_off = [], [], [], [], []
        byte_start = np.zeros(P + 1, np.int64)
        for p in range(P):
            pbytes = 0
            offs_p = [np.zeros(1, np.int32)]
            for s in segs:
                rs = s["row_start"]
                a, b = int(rs[p]), int(rs[p + 1])
                sc = s["columns"][c]
                if has_valid:
                    parts_valid.append(_bits(sc["validity"], a, b - a))
                if t == T.BOOL:
                    parts_bool.append(_bits(sc["values"], a, b - a))
                elif T.is_binary(t):
                    o = sc["offsets"][a + p: b + p + 1]
                    bs = sc["byte_start"]
                    parts_data.append(sc["data"][int(bs[p]): int(bs[p]) + int(o[-1])])
                    offs_p.append(o[1:] + pbytes)
                    pbytes += int(o[-1])
                elif t != T.NA:
                    w = T.width(t)
                    parts_vals.append(sc["values"][a * w: b * w])
            if T.is_binary(t):
                parts_off.append(np.concatenate(offs_p).astype(np.int32))
                byte_start[p + 1] = byte_start[p] + pbytes
        nbm = (N + 63) // 64 * 64
        if has_valid:
            v = np.ones(nbm, bool)
            v[:N] = np.concatenate(parts_valid) if parts_valid else np.zeros(0, bool)
            col["validity"] = np.packbits(v, bitorder="little")
        else:
            col["validity"] = None
        if t == T.BOOL:
            v = np.ones(nbm, bool)
            v[:N] = np.concatenate(parts_bool) if parts_bool else np.zeros(0, bool)
            col["values"] = np.packbits(v, bitorder="little")
        elif T.is_binary(t):
            col["offsets"] = np.concatenate(parts_off) if parts_off else np.zeros(0, np.int32)
            col["data"] = np.concatenate(parts_data) if parts_data else np.zeros(0, np.uint8)
            col["byte_start"] = byte_start
        elif t != T.NA:
            col["values"] = np.concatenate(parts_vals) if parts_vals else np.zeros(0, np.uint8)
        cols.append(col)
    return {"row_start": row_start, "num_rows": N, "columns": cols}
