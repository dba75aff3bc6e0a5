"""Fragment server (the serving form) over the C ABI: continuous batching of
aggregate, join-then-aggregate and shuffle-split fragments on one device, with
admission by HBM budget and resident hash relations.  The loop itself is native
(gz_server_*); this module builds the plans and fragment lists and reads back
results and per-fragment latencies."""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib as L
from .aggregate import _NP, _actions_c, parse_action
from .hip import to_host
from .splitter import Splitter

AGGREGATE, JOIN_AGGREGATE, SPLIT = 0, 1, 2
JOIN_TYPES = {"inner": 0, "anti": 2, "semi": 3}


def _keep(store, obj):
    store.append(obj)
    return obj


class FragmentServer:
    def __init__(self, device: int = 0, hbm_budget: int = 200 << 30, slots: int = 4,
                 agg: Optional[Tuple[Sequence[int], Sequence]] = None,
                 split: Optional[Tuple[Sequence[int], int, Sequence[int]]] = None,
                 probe: Optional[Tuple[Sequence[int], int, str, Sequence[Tuple[int, int]]]] = None,
                 join_agg: Optional[Tuple[Sequence[int], Sequence]] = None):
        """agg / join_agg: (column types, [(reference action name, inputs)]);
        split: (column types, partitions, key columns); probe: (probe types, probe
        key column, join type, result columns [(side, index)])."""
        self._store = []
        self.agg_types = list(agg[0]) if agg else None
        self.join_agg_types = list(join_agg[0]) if join_agg else None
        self.split_types = list(split[0]) if split else None
        self.split_P = split[1] if split else 0
        ap = self._agg_plan(agg)
        jp = self._agg_plan(join_agg)
        sp = pp = None
        if split:
            sp = _keep(self._store, L.GzSplitPlan(len(split[0]), _keep(self._store, L.i32_array(split[0])), split[1],
                                                  _keep(self._store, L.i32_array(split[2])), len(split[2])))
        if probe:
            sides = [s for s, _ in probe[3]]
            idx = [i for _, i in probe[3]]
            pp = _keep(self._store, L.GzProbePlan(len(probe[0]), _keep(self._store, L.i32_array(probe[0])), probe[1],
                                                  JOIN_TYPES[probe[2]], len(sides),
                                                  _keep(self._store, L.i32_array(sides)),
                                                  _keep(self._store, L.i32_array(idx))))
        self._h = ctypes.c_void_p()
        ref = lambda x: ctypes.byref(x) if x is not None else None
        L.check(L.lib().gz_server_create(device, hbm_budget, slots, ref(ap), ref(sp), ref(pp), ref(jp),
                                         ctypes.byref(self._h)))
        self.relations = []

    def _agg_plan(self, agg):
        if not agg:
            return None
        acts = [parse_action(n, i) if isinstance(n, str) else (n, *i) for n, i in agg[1]]
        return _keep(self._store, L.GzAggPlan(len(agg[0]), _keep(self._store, L.i32_array(agg[0])), len(acts),
                                              _keep(self._store, _actions_c(acts))))

    def add_relation(self, rel) -> int:
        out = ctypes.c_int32()
        L.check(L.lib().gz_server_add_relation(self._h, rel.handle, ctypes.byref(out)))
        self.relations.append(rel)
        return out.value

    def run(self, fragments: Sequence[Tuple[int, int, object, int, int, float]], keep: bool = False) -> np.ndarray:
        """fragments: (kind, relation, batch array, first batch, batch count,
        arrival seconds), sorted by arrival.  Returns a structured array of
        per-fragment times and sizes."""
        n = len(fragments)
        arr = (L.GzFragment * max(1, n))()
        size = ctypes.sizeof(L.GzBatch)
        for i, (kind, rel, barr, b0, nb, t) in enumerate(fragments):
            arr[i].kind = kind
            arr[i].relation = rel
            arr[i].num_batches = nb
            arr[i].batches = ctypes.cast(ctypes.addressof(barr) + b0 * size, ctypes.POINTER(L.GzBatch))
            arr[i].arrival_s = t
        stats = (L.GzFragmentStats * max(1, n))()
        L.check(L.lib().gz_server_run(self._h, n, arr, int(keep), stats))
        dt = np.dtype([("arrival_s", "f8"), ("admitted_s", "f8"), ("done_s", "f8"), ("rows", "i8"),
                       ("out_rows", "i8"), ("status", "i4"), ("slot", "i4")])
        return np.frombuffer(bytes(stats), dtype=dt, count=n).copy()

    def agg_result(self, i: int, types: Sequence[int], nres: int) -> List:
        """Result columns (type, values, valid) of aggregate / join fragment i."""
        h = ctypes.c_void_p()
        L.check(L.lib().gz_server_fragment_agg(self._h, i, ctypes.byref(h)))
        g = ctypes.c_int64()
        L.check(L.lib().gz_agg_num_groups(h, ctypes.byref(g)))
        out = []
        for c in range(L.lib().gz_agg_num_result_columns(h)):
            t = L.lib().gz_agg_result_type(h, c)
            vals, valid = ctypes.c_void_p(), ctypes.c_void_p()
            L.check(L.lib().gz_agg_result_column(h, c, ctypes.byref(vals), ctypes.byref(valid)))
            dt = np.dtype(_NP[t])
            n = g.value
            v = to_host(vals.value, n * dt.itemsize, dt) if n else np.zeros(0, dt)
            ok = to_host(valid.value, n, np.uint8).astype(bool) if n else np.zeros(0, bool)
            out.append((t, v, ok))
        return out

    def split_result(self, i: int):
        """Partitioned layout of split fragment i (as Splitter.read_partitions)."""
        h = ctypes.c_void_p()
        L.check(L.lib().gz_server_fragment_split(self._h, i, ctypes.byref(h)))
        sp = Splitter(None, h, self.split_types, self.split_P, "hash")
        try:
            return sp.read_partitions()
        finally:
            sp._h = ctypes.c_void_p()  # owned by the server

    def close(self):
        if self._h:
            L.lib().gz_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def latency_summary(stats: np.ndarray) -> dict:
    lat = stats["done_s"] - stats["arrival_s"]
    span = float(stats["done_s"].max() - stats["arrival_s"].min()) if len(stats) else 0.0
    return {"fragments": int(len(stats)), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "rows": int(stats["rows"].sum()),
            "rows_per_s": float(stats["rows"].sum() / span) if span > 0 else 0.0, "span_s": span}
