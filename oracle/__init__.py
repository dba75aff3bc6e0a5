"""Python wrapper of the CPU oracle (libgz_oracle.so).

TEST INFRASTRUCTURE ONLY: this package is the checker for the HIP path
(``tests/``, ``__graft_entry__.smoke``) and the CPU baseline leg of ``bench.py``.
The product (``gazelle_plugin_amd``) never imports it.

Each function restates a reference algorithm; see the C sources for the file:line
each one follows:
  split.c  -- Splitter hash partitioning + DoSplit   (cpp/src/shuffle/splitter.cc)
  agg.c    -- HashAggregateResultIterator + actions  (ext/hash_aggregate_kernel.cc, ext/actions_impl.cc)
  join.c   -- HashRelation / unsafeHashMap + probes  (codegen/common/hash_relation.h,
              third_party/row_wise_memory/hashMap.h, ext/conditioned_probe_kernel.cc)
  lz4.c    -- LZ4_FRAME decoding of shuffle buffers   (jni/jni_common.h:471-492, lz4 frame format)

Pins (what checks the oracle itself) are listed in oracle/README.md.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Dict, List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libgz_oracle.so")
_lib = None

_T_NA, _T_BOOL, _T_STRING, _T_BINARY = 0, 1, 13, 14
_WIDTH = {2: 1, 3: 1, 4: 2, 5: 2, 6: 4, 7: 4, 11: 4, 16: 4, 8: 8, 9: 8, 12: 8, 17: 8, 18: 8, 23: 16}


class _Batch(ctypes.Structure):
    _fields_ = [("num_rows", ctypes.c_int64), ("bufs", ctypes.POINTER(ctypes.c_void_p))]


def build() -> str:
    subprocess.check_call(["make", "-s", "-C", _HERE])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.gzo_hash_int.restype = ctypes.c_int32
        L.gzo_hash_int.argtypes = [ctypes.c_int32, ctypes.c_int32]
        L.gzo_hash_long.restype = ctypes.c_int32
        L.gzo_hash_long.argtypes = [ctypes.c_int64, ctypes.c_int32]
        L.gzo_hash_bytes.restype = ctypes.c_int32
        L.gzo_hash_bytes.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32]
        L.gzo_hash32_numeric.restype = ctypes.c_int32
        L.gzo_hash32_numeric.argtypes = [ctypes.c_double]
        L.gzo_split_pids.restype = ctypes.c_int
        L.gzo_splitter_make.restype = ctypes.c_void_p
        L.gzo_splitter_rows.restype = ctypes.c_int64
        L.gzo_splitter_rows.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.gzo_splitter_split.argtypes = [ctypes.c_void_p, ctypes.POINTER(_Batch)]
        L.gzo_splitter_free.argtypes = [ctypes.c_void_p]
        L.gzo_agg_make.restype = ctypes.c_void_p
        L.gzo_agg_process.argtypes = [ctypes.c_void_p, ctypes.POINTER(_Batch)]
        L.gzo_agg_num_groups.restype = ctypes.c_int64
        L.gzo_agg_num_groups.argtypes = [ctypes.c_void_p]
        L.gzo_agg_num_results.argtypes = [ctypes.c_void_p]
        L.gzo_agg_result_type.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.gzo_agg_result.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.gzo_agg_result_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.gzo_agg_result_bytes.restype = ctypes.c_int64
        L.gzo_agg_result_binary.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p]
        L.gzo_agg_free.argtypes = [ctypes.c_void_p]
        L.gzo_rel_make.restype = ctypes.c_void_p
        L.gzo_rel_make.argtypes = [ctypes.c_int32, ctypes.c_int32]
        L.gzo_rel_append.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        L.gzo_rel_finish.argtypes = [ctypes.c_void_p]
        L.gzo_rel_num_keys.restype = ctypes.c_int64
        L.gzo_rel_num_keys.argtypes = [ctypes.c_void_p]
        L.gzo_rel_has_null.argtypes = [ctypes.c_void_p]
        L.gzo_lz4f_decompress.restype = ctypes.c_int64
        L.gzo_lz4f_decompress.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64]
        L.gzo_probe.restype = ctypes.c_int64
        L.gzo_probe.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        L.gzo_rel_free.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class _CBatches:
    """Keeps the ctypes views of a list of HostBatch-like objects alive."""

    def __init__(self, batches):
        self.keep = []
        self.arr = (_Batch * max(1, len(batches)))()
        for i, b in enumerate(batches):
            slots = []
            for c in b.columns:
                for a in (c.validity, c.offsets, c.values):
                    if a is not None:
                        a = np.ascontiguousarray(a)
                        self.keep.append(a)
                        slots.append(a.ctypes.data)
                    else:
                        slots.append(None)
            bufs = (ctypes.c_void_p * max(1, len(slots)))(*slots)
            self.keep.append(bufs)
            self.arr[i].num_rows = b.num_rows
            self.arr[i].bufs = ctypes.cast(bufs, ctypes.POINTER(ctypes.c_void_p))


# ------------------------------------------------------------------ hashing

def hash_int(v: int, seed: int = 42) -> int:
    return lib().gzo_hash_int(int(np.int32(v)), seed)


def hash_long(v: int, seed: int = 42) -> int:
    return lib().gzo_hash_long(int(np.int64(v)), seed)


def hash_bytes(b: bytes, seed: int = 42) -> int:
    return lib().gzo_hash_bytes(b, len(b), seed)


def pmod(h: int, p: int) -> int:
    r = int(np.fmod(np.int64(h), p))
    return r + p if r < 0 else r


# ------------------------------------------------------------------ split

MODES = {"hash": 0, "rr": 1, "range": 2, "single": 3}


def split_pids(batches, types: Sequence[int], P: int, keys: Sequence[int] = (), mode: str = "hash",
               seed: int = 42) -> np.ndarray:
    L = lib()
    out = []
    rr = ctypes.c_int32(0)
    t = (ctypes.c_int32 * len(types))(*types)
    k = (ctypes.c_int32 * max(1, len(keys)))(*keys)
    for b in batches:
        cb = _CBatches([b])
        pids = np.zeros(max(1, b.num_rows), dtype=np.int32)
        rc = L.gzo_split_pids(ctypes.c_int(MODES[mode]), ctypes.byref(cb.arr[0]), len(types), t, k,
                              len(keys), ctypes.c_int(P if mode != "single" else 1),
                              ctypes.c_int32(seed), ctypes.byref(rr), _ptr(pids))
        if rc != 0:
            raise ValueError(f"oracle split_pids failed with status {rc}")
        out.append(pids[: b.num_rows])
    return np.concatenate(out) if out else np.zeros(0, np.int32)


def split(batches, types: Sequence[int], P: int, keys: Sequence[int] = (), mode: str = "hash",
          seed: int = 42, pids: Optional[np.ndarray] = None) -> Dict:
    """Concatenated partition layout of the reference split of ``batches``.

    For mode "range" column 0 carries the pids and is dropped from the output,
    as FallbackRangeSplitter does.
    """
    L = lib()
    if pids is None:
        pids = split_pids(batches, types, P, keys, mode, seed)
    if mode == "single":
        P = 1
    if mode == "range":
        batches = [type(b)(b.num_rows, b.columns[1:]) for b in batches]
        types = list(types[1:])
    ncols = len(types)
    nbin = sum(1 for t in types if t in (_T_STRING, _T_BINARY))
    N = int(sum(b.num_rows for b in batches))
    cb = _CBatches(batches)
    tarr = (ctypes.c_int32 * max(1, ncols))(*types)
    row_start = np.zeros(P + 1, dtype=np.int64)
    byte_start = np.zeros(max(1, nbin) * (P + 1), dtype=np.int64)
    pids = np.ascontiguousarray(pids, dtype=np.int32)
    L.gzo_split_layout(cb.arr, len(batches), ncols, tarr, P, _ptr(pids), _ptr(row_start), _ptr(byte_start))
    values, validity, offsets, data = [], [], [], []
    j = 0
    bs_list = []
    for c, t in enumerate(types):
        has_valid = any(b.columns[c].validity is not None for b in batches)
        nwords = (N + 63) // 64
        validity.append(np.full(nwords * 8, 0xFF, np.uint8) if has_valid and t != _T_NA else None)
        if t in (_T_STRING, _T_BINARY):
            bs = byte_start[j * (P + 1):(j + 1) * (P + 1)].copy()
            bs_list.append(bs)
            offsets.append(np.zeros(N + P, np.int32))
            data.append(np.zeros(max(1, int(bs[-1])), np.uint8))
            values.append(None)
            j += 1
        else:
            bs_list.append(None)
            offsets.append(None)
            data.append(None)
            if t == _T_BOOL:
                values.append(np.full(nwords * 8, 0xFF, np.uint8))
            elif t == _T_NA:
                values.append(None)
            else:
                values.append(np.zeros(max(1, N) * _WIDTH[t], np.uint8))
    vp = (ctypes.c_void_p * max(1, ncols))(*[v.ctypes.data if v is not None else None for v in values])
    vb = (ctypes.c_void_p * max(1, ncols))(*[v.ctypes.data if v is not None else None for v in validity])
    op = (ctypes.c_void_p * max(1, ncols))(*[v.ctypes.data if v is not None else None for v in offsets])
    dp = (ctypes.c_void_p * max(1, ncols))(*[v.ctypes.data if v is not None else None for v in data])
    L.gzo_split_scatter(cb.arr, len(batches), ncols, tarr, P, _ptr(pids), _ptr(row_start),
                        _ptr(byte_start), vp, vb, op, dp)
    cols = []
    for c, t in enumerate(types):
        col = {"type": t, "validity": validity[c]}
        if t in (_T_STRING, _T_BINARY):
            col["offsets"] = offsets[c]
            col["data"] = data[c][: int(bs_list[c][-1])]
            col["byte_start"] = bs_list[c]
        elif values[c] is not None:
            col["values"] = values[c] if t == _T_BOOL else values[c][: N * _WIDTH[t]]
        cols.append(col)
    return {"P": P, "num_rows": N, "row_start": row_start, "pids": pids, "columns": cols}


class CpuSplitter:
    """The reference per-batch split loop (for timing the CPU baseline)."""

    def __init__(self, types: Sequence[int], P: int, keys: Sequence[int], seed: int = 42):
        L = lib()
        L.gzo_splitter_make.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_int, ctypes.c_int32]
        self._t = (ctypes.c_int32 * len(types))(*types)
        self._k = (ctypes.c_int32 * max(1, len(keys)))(*keys)
        self.P = P
        self.h = L.gzo_splitter_make(len(types), self._t, P, self._k, len(keys), seed)

    def split(self, cbatches: "_CBatches", i: int):
        lib().gzo_splitter_split(self.h, ctypes.byref(cbatches.arr[i]))

    def split_many(self, cbatches: "_CBatches", i0: int, nb: int):
        """Batches i0 .. i0 + nb in one C loop (the GIL is released meanwhile)."""
        L = lib()
        L.gzo_splitter_split_many.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.gzo_splitter_split_many(self.h, ctypes.byref(cbatches.arr, i0 * ctypes.sizeof(_Batch)), nb)

    def rows(self) -> np.ndarray:
        return np.array([lib().gzo_splitter_rows(self.h, p) for p in range(self.P)], np.int64)

    def close(self):
        if self.h:
            lib().gzo_splitter_free(self.h)
            self.h = None


# ------------------------------------------------------------------ aggregate

class CpuAggregate:
    """The reference hash aggregate over whole batch slices (for timing the CPU
    baseline): actions are (kind, in0, in1, nan_check) tuples."""

    def __init__(self, types: Sequence[int], actions: Sequence[Sequence[int]]):
        L = lib()
        L.gzo_agg_make.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.gzo_agg_process_many.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        na = len(actions)
        self._keep = [(ctypes.c_int32 * na)(*[a[j] for a in actions]) for j in range(4)]
        self._t = (ctypes.c_int32 * len(types))(*types)
        self.h = L.gzo_agg_make(len(types), self._t, na, *self._keep)

    def process_many(self, cbatches: "_CBatches", i0: int, nb: int):
        rc = lib().gzo_agg_process_many(self.h, ctypes.byref(cbatches.arr, i0 * ctypes.sizeof(_Batch)), nb)
        if rc:
            raise ValueError(f"oracle aggregate failed with status {rc}")

    def num_groups(self) -> int:
        return lib().gzo_agg_num_groups(self.h)

    def close(self):
        if self.h:
            lib().gzo_agg_free(self.h)
            self.h = None


class CpuRelation:
    """The reference hash relation over one int64 key column, built once and
    probed from any number of threads (the probe only reads it)."""

    def __init__(self, key_type: int = 9, builder_type: int = 1):
        self.h = lib().gzo_rel_make(key_type, builder_type)
        self._keep = []

    def append(self, keys: np.ndarray):
        keys = np.ascontiguousarray(keys)
        self._keep.append(keys)
        lib().gzo_rel_append(self.h, _ptr(keys), None, len(keys))

    def finish(self):
        lib().gzo_rel_finish(self.h)

    def probe(self, keys: np.ndarray, join_type: int, out_probe: np.ndarray, out_build: np.ndarray) -> int:
        return lib().gzo_probe(self.h, join_type, 0, _ptr(keys), None, len(keys), _ptr(out_probe), _ptr(out_build))

    def close(self):
        if self.h:
            lib().gzo_rel_free(self.h)
            self.h = None

class ResultColumn(tuple):
    """(type, values, valid) of one result column; ``approx`` is True for a
    floating sum or average (its summation order is free, so a check allows a
    relative tolerance) and False for every other column, whose bits are the
    reference's (min / max keep one input value, NaN payload and zero sign
    included; group keys echo an input value)."""

    approx = False


_A_SUM, _A_AVG, _A_SUM_COUNT, _A_AVG_BY_COUNT = 3, 4, 7, 8


def _approx_columns(actions) -> List[bool]:
    out = []
    for a in actions:
        k = a[0]
        if k == _A_SUM_COUNT:
            out += [True, False]
        else:
            out.append(k in (_A_SUM, _A_AVG, _A_AVG_BY_COUNT))
    return out


def _result(t, vals, valid, approx: bool) -> ResultColumn:
    r = ResultColumn((t, vals, valid))
    r.approx = approx
    return r


def aggregate(batches, types: Sequence[int], actions: Sequence[Sequence[int]]) -> List:
    """actions: (kind, in0, in1, nan_check) tuples.  Returns [(type, values, valid)]
    per result column (``ResultColumn``), groups in first-seen order."""
    approx = _approx_columns(actions)
    L = lib()
    L.gzo_agg_make.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    na = len(actions)
    kinds = (ctypes.c_int32 * na)(*[a[0] for a in actions])
    in0 = (ctypes.c_int32 * na)(*[a[1] for a in actions])
    in1 = (ctypes.c_int32 * na)(*[a[2] for a in actions])
    nan = (ctypes.c_int32 * na)(*[a[3] for a in actions])
    tarr = (ctypes.c_int32 * len(types))(*types)
    h = L.gzo_agg_make(len(types), tarr, na, kinds, in0, in1, nan)
    try:
        for b in batches:
            cb = _CBatches([b])
            L.gzo_agg_process(h, ctypes.byref(cb.arr[0]))
        g = L.gzo_agg_num_groups(h)
        out = []
        for c in range(L.gzo_agg_num_results(h)):
            t = L.gzo_agg_result_type(h, c)
            if t in (13, 14):  # a utf8 key: bytes objects (None-free; see valid)
                nbytes = L.gzo_agg_result_bytes(h, c)
                offs = np.zeros(g + 1, np.int32)
                data = np.zeros(max(1, nbytes), np.uint8)
                valid = np.zeros(max(1, g), np.uint8)
                L.gzo_agg_result_binary(h, c, _ptr(offs), _ptr(data), _ptr(valid))
                vals = np.array([data[offs[i]:offs[i + 1]].tobytes() for i in range(g)] + [None], object)[:g]
                out.append(_result(t, vals, valid[:g].astype(bool), False))
                continue
            vals = np.zeros(max(1, g) * _WIDTH.get(t, 8), np.uint8)
            valid = np.zeros(max(1, g), np.uint8)
            L.gzo_agg_result(h, c, _ptr(vals), _ptr(valid))
            dt = {2: np.uint8, 3: np.int8, 4: np.uint16, 5: np.int16, 6: np.uint32, 7: np.int32,
                  8: np.uint64, 9: np.int64, 11: np.float32, 12: np.float64, 16: np.int32,
                  17: np.int64, 18: np.int64}[t]
            out.append(_result(t, vals[: g * _WIDTH.get(t, 8)].view(dt).copy(), valid[:g].astype(bool),
                               approx[c] and t in (11, 12)))
        return out
    finally:
        L.gzo_agg_free(h)


# ------------------------------------------------------------------ join

def normalize_nan_zero(values: np.ndarray) -> np.ndarray:
    """Spark's join-key "normalize" on a float/double array, as the reference's
    normalize_nan_zero (precompile/gandiva.h:214-222): NaN -> NaN, |x| < 1e-7
    -> 0.0, computed in double and stored back in the array's type."""
    d = np.asarray(values).astype(np.float64)
    with np.errstate(invalid="ignore"):
        out = np.where(np.isnan(d), np.nan, np.where(np.abs(d) < 0.0000001, 0.0, d))
    return out.astype(np.asarray(values).dtype)


def join(build_keys: Sequence, probe_keys, key_type: int, join_type: int, null_aware: bool = False,
         builder_type: int = 1, normalize: bool = False):
    """build_keys: list of (values ndarray, validity bitmap or None) per build batch.
    probe_keys: list of (values, validity, n) per probe batch.
    Returns per probe batch (probe_rows, build_rows) for inner, probe_rows otherwise.
    Build rows are global indices (batch order, then row order).  normalize: the
    float/double key values go through normalize_nan_zero on both sides first
    (the key expressions of the reference's join, which then compares the typed
    keys as usual: a NaN key still matches nothing)."""
    if normalize:
        build_keys = [(normalize_nan_zero(v), valid, n) for v, valid, n in build_keys]
        probe_keys = [(normalize_nan_zero(v), valid, n) for v, valid, n in probe_keys]
    L = lib()
    r = L.gzo_rel_make(key_type, builder_type)
    try:
        keep = []
        for vals, valid, n in build_keys:
            vals = np.ascontiguousarray(vals)
            keep.append(vals)
            L.gzo_rel_append(r, _ptr(vals), _ptr(valid), n)
        L.gzo_rel_finish(r)
        outs = []
        for vals, valid, n in probe_keys:
            vals = np.ascontiguousarray(vals)
            cnt = L.gzo_probe(r, join_type, int(null_aware), _ptr(vals), _ptr(valid), n, None, None)
            pr = np.zeros(max(1, cnt), np.int64)
            br = np.full(max(1, cnt), -1, np.int64)
            L.gzo_probe(r, join_type, int(null_aware), _ptr(vals), _ptr(valid), n, _ptr(pr), _ptr(br))
            outs.append((pr[:cnt], br[:cnt]) if join_type == 0 else pr[:cnt])
        return outs
    finally:
        L.gzo_rel_free(r)


def _key_tuples(batch, cols, normalize=None) -> list:
    """Per row of a HostBatch: None if any of the key columns is null, else the
    tuple of the keys' raw bits (a utf8 key's bytes); normalize[j] puts key j
    through normalize_nan_zero first."""
    n = batch.num_rows
    parts = []
    for j, c in enumerate(cols):
        col = batch.columns[c]
        ok = (np.ones(n, bool) if col.validity is None else
              np.unpackbits(np.asarray(col.validity, np.uint8), bitorder="little")[:n].astype(bool))
        if col.type_id in (13, 14):
            vals = [bytes(col.values[col.offsets[r]:col.offsets[r + 1]]) for r in range(n)]
        else:
            v = col.values[:n]
            if normalize is not None and normalize[j]:
                v = normalize_nan_zero(v)
            raw = v.view(np.dtype(f"u{col.values.dtype.itemsize}"))
            vals = [int(v) for v in raw]
        parts.append((vals, ok))
    return [None if not all(ok[r] for _, ok in parts) else tuple(v[r] for v, _ in parts) for r in range(n)]


def join_multi(build, build_key_cols, probe, probe_key_cols, join_type: int, builder_type: int = 1,
               null_aware: bool = False, normalize=None):
    """Hash join on several key columns (or one utf8 key).  The reference keys
    such relations by UnsafeRow bytes: a build row with a null key is never
    inserted (hash_relation.h:176-198), and a probe row matches the build rows
    whose key bytes are equal (conditioned_probe_kernel.cc:609-640), so a probe
    row with a null key matches nothing.  Restated as the one-key join over
    first-seen ids of the build side's key tuples: a build row with a null key
    gets a null id, so does a probe row with a null key, and a tuple the build
    side lacks gets id -1.  The UnsafeRow anti probe never reads the null-aware
    flag (:957-982), so it applies to one utf8 key only, with the one-key rules
    (a null build key is the relation's null key, InsertNull, hash_relation.h:
    244-270).  normalize[j]: key j goes through normalize_nan_zero on both
    sides.  Same return as join()."""
    def bitmap(ok):
        bits = np.ones((len(ok) + 63) // 64 * 64, np.uint8)
        bits[:len(ok)] = ok
        return np.packbits(bits, bitorder="little")

    ids = {}
    bk, pk = [], []
    for b in build:
        t = _key_tuples(b, build_key_cols, normalize)
        vals = np.array([ids.setdefault(k, len(ids)) if k is not None else 0 for k in t], np.int64)
        bk.append((vals, bitmap(np.array([k is not None for k in t], np.uint8)), b.num_rows))
    for b in probe:
        t = _key_tuples(b, probe_key_cols, normalize)
        got = [ids.get(k, -1) if k is not None else None for k in t]
        vals = np.array([g if g is not None else 0 for g in got], np.int64)
        pk.append((vals, bitmap(np.array([g is not None for g in got], np.uint8)), b.num_rows))
    return join(bk, pk, 9, join_type, null_aware and len(build_key_cols) == 1, builder_type)


# ------------------------------------------------------------------ lz4

def lz4f_decompress(frame, capacity: int) -> np.ndarray:
    """LZ4 frame(s) -> bytes (at most ``capacity``); ValueError on malformed input."""
    src = np.frombuffer(bytes(frame), np.uint8) if not isinstance(frame, np.ndarray) else \
        np.ascontiguousarray(frame, np.uint8)
    out = np.zeros(max(1, capacity), np.uint8)
    n = lib().gzo_lz4f_decompress(_ptr(src) if len(src) else None, len(src), _ptr(out), capacity)
    if n < 0:
        raise ValueError("malformed LZ4 frame")
    return out[:n]


def ipc_decompress_buffer(buf) -> np.ndarray:
    """One compressed IPC body buffer (int64 uncompressed length, then an LZ4
    frame; length -1: stored as is), as DecompressBuffer (jni_common.h:471-492)
    reads it."""
    b = np.frombuffer(bytes(buf), np.uint8)
    if len(b) == 0:
        return b
    if len(b) < 8:
        raise ValueError("compressed buffer shorter than its length prefix")
    n = int(b[:8].view(np.int64)[0])
    if n == -1:
        return b[8:].copy()
    out = lz4f_decompress(b[8:], n)
    if len(out) != n:
        raise ValueError(f"decompressed {len(out)} bytes, expected {n}")
    return out


# ------------------------------------------------------------------ partition buffers

_BITWIDTH_BYTES = {1: 0, 2: 1, 3: 1, 4: 2, 5: 2, 6: 4, 7: 4, 8: 8, 9: 8, 11: 4, 12: 8, 16: 4, 17: 8, 18: 8, 23: 16}


def split_buffer_policy(batches, types: Sequence[int], pids_per_batch, P: int, buffer_size: int = 4096,
                        offheap_per_task: int = 0, prefer_spill: bool = True) -> Dict:
    """The reference's partition buffers over a task's input batches, batch by
    batch as Splitter::DoSplit runs them (cpp/src/shuffle/splitter.cc:855-938):
    CalculateSplitBatchSize (:800-853), AllocatePartitionBuffers' validity rule
    (:610-713), SplitValidityBuffer's lazy validity (:1179-1207),
    CacheRecordBatch's cuts and batch_nbytes (:438-608), Stop (:395-437).

    Returns {"batch_rows": the last CalculateSplitBatchSize value, "cuts": rows
    of each cached record batch per partition, "raw": RawPartitionLengths}."""
    ncol = len(types)
    fixed = [c for c in range(ncol) if types[c] not in (_T_NA, _T_STRING, _T_BINARY)]
    binary = [c for c in range(ncol) if types[c] in (_T_STRING, _T_BINARY)]
    array_idx = fixed + binary  # slot -> schema column
    fixed_cnt = len(fixed)
    ns = len(array_idx)
    input_has_null = [False] * ns
    emp = [0] * len(binary)
    size = [0] * P
    base = [0] * P
    valid = [[False] * ns for _ in range(P)]
    bytes_in = [[0] * len(binary) for _ in range(P)]  # utf8 bytes in the partition buffer
    cuts = [[] for _ in range(P)]
    raw = [0] * P
    last_bs = buffer_size

    def null_count(b, c):
        vb = b.columns[c].validity
        if vb is None:
            return 0
        bits = np.unpackbits(np.asarray(vb, np.uint8), bitorder="little")[: b.num_rows]
        return int(b.num_rows - bits.sum())

    def calc(b):
        nonlocal last_bs
        n = b.num_rows
        for j, c in enumerate(binary):
            if emp[j] == 0:
                off = b.columns[c].offsets
                length = int(off[n]) - int(off[0])
                emp[j] = length // n if length % n == 0 else length // n + 1
        spr = sum(emp)
        for c in array_idx:
            if c < fixed_cnt:  # the schema index, as the reference compares it
                spr += _BITWIDTH_BYTES.get(types[c], 0)
        pre = (offheap_per_task // spr // P) >> 2 if offheap_per_task > 0 and spr > 0 else buffer_size
        last_bs = min(pre, buffer_size)
        return last_bs

    def allocate(p, new):
        size[p] = new
        valid[p] = [input_has_null[k] for k in range(ns)]

    def cache(p, reset):
        n = base[p]
        if n > 0:
            total = 0
            for k, c in enumerate(array_idx):
                t = types[c]
                if valid[p][k]:
                    total += (n + 7) // 8
                if t in (_T_STRING, _T_BINARY):
                    total += (n + 1) * 4 + bytes_in[p][binary.index(c)]
                elif t == _T_BOOL:
                    total += (n + 7) // 8
                else:
                    total += n * _WIDTH[t]
            raw[p] += total
            cuts[p].append(n)
        base[p] = 0
        bytes_in[p] = [0] * len(binary)
        if reset:
            valid[p] = [False] * ns

    for b, pids in zip(batches, pids_per_batch):
        if b.num_rows == 0:
            continue
        pids = np.asarray(pids)
        cnt = np.bincount(pids, minlength=P)
        nulls = {c: null_count(b, c) for c in array_idx}
        for k, c in enumerate(array_idx):
            if c < fixed_cnt and not input_has_null[k] and nulls[c] != 0:
                input_has_null[k] = True
        for p in range(P):
            if cnt[p] == 0:
                continue
            if size[p] == 0:
                allocate(p, max(calc(b), int(cnt[p])))
            elif base[p] + cnt[p] > size[p]:
                new = max(calc(b), int(cnt[p]))
                if prefer_spill:
                    if new > size[p]:
                        cache(p, True)
                        allocate(p, new)
                    else:
                        cache(p, False)
                else:
                    cache(p, True)
                    allocate(p, new)
        for k, c in enumerate(array_idx):
            if nulls[c] > 0:
                for p in np.nonzero(cnt)[0]:
                    valid[p][k] = True
        for j, c in enumerate(binary):
            off = np.asarray(b.columns[c].offsets, np.int64)
            lens = off[1:b.num_rows + 1] - off[:b.num_rows]
            per = np.bincount(pids, weights=lens, minlength=P)
            for p in range(P):
                bytes_in[p][j] += int(per[p])
        for p in range(P):
            base[p] += int(cnt[p])
    for p in range(P):
        cache(p, True)
    return {"batch_rows": last_bs, "cuts": cuts, "raw": raw}
