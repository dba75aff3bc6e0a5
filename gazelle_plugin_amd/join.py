"""Python mirror of the reference hash join over the C ABI.

``HashRelationKernel`` mirrors the build (ext/hash_relation_kernel.cc:50-344):
``Evaluate`` caches a build batch, ``Finish`` builds the table (FinishInternal).
``ConditionedProbeKernel`` mirrors the probe (ext/conditioned_probe_kernel.cc):
``Make(relation, ..., join_type, is_null_aware_anti_join, result)``, then
``Process(batch)`` returns the output record batch of that probe batch, rows in
probe order with each probe row's matches in build insertion order.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib as L
from . import types as T
from .batch import DeviceBatch
from .context import Context
from .hip import to_host
from .splitter import make_batch_array

JOIN_TYPES = {"inner": 0, "anti": 2, "semi": 3}

_NP = {T.UINT8: np.uint8, T.INT8: np.int8, T.UINT16: np.uint16, T.INT16: np.int16,
       T.UINT32: np.uint32, T.INT32: np.int32, T.UINT64: np.uint64, T.INT64: np.int64,
       T.FLOAT: np.float32, T.DOUBLE: np.float64, T.DATE32: np.int32, T.DATE64: np.int64,
       T.TIMESTAMP: np.int64, T.BOOL: np.uint8}


class HashRelation:
    """The build side: a resident table in HBM (an opaque handle)."""

    def __init__(self, ctx: Context, types: Sequence[int], key_col, builder_type: int = 1, normalize=None):
        """key_col: a column index, or a list of them (several key columns).
        normalize: per key, apply Spark's "normalize" to a float/double key
        (gz_relation_normalize_keys); a bool for one key."""
        self.ctx = ctx
        self.types = list(types)
        self.key_col = key_col
        self._h = ctypes.c_void_p()
        if isinstance(key_col, (list, tuple)):  # several key columns (UnsafeRow keys)
            L.check(L.lib().gz_relation_make_keys(ctx.handle, len(types), L.i32_array(types), len(key_col),
                                                  L.i32_array(list(key_col)), builder_type, ctypes.byref(self._h)))
        else:
            L.check(L.lib().gz_relation_make(ctx.handle, len(types), L.i32_array(types), key_col, builder_type,
                                             ctypes.byref(self._h)))
        if normalize is not None:
            flags = [int(bool(x)) for x in (normalize if isinstance(normalize, (list, tuple)) else [normalize])]
            if any(flags):
                L.check(L.lib().gz_relation_normalize_keys(self._h, len(flags), L.i32_array(flags)))
        self.finished = False

    def Evaluate(self, batch: DeviceBatch):
        self.ctx.hold(batch)
        L.check(L.lib().gz_relation_append(self._h, batch.num_rows, batch.addrs, batch.sizes, batch.num_bufs))

    def EvaluateBatches(self, batches: Sequence[DeviceBatch], arr=None):
        """arr: a prebuilt gz_batch array of the batches (make_batch_array)."""
        self.ctx.hold(list(batches))
        if arr is None:
            arr = make_batch_array(batches)
        L.check(L.lib().gz_relation_append_batches(self._h, len(batches), arr))

    def Finish(self):
        L.check(L.lib().gz_relation_finish(self._h))
        self.finished = True
        return self

    def num_keys(self) -> int:
        n = ctypes.c_int64()
        L.check(L.lib().gz_relation_num_keys(self._h, ctypes.byref(n)))
        return n.value

    def has_null_key(self) -> bool:
        n = ctypes.c_int32()
        L.check(L.lib().gz_relation_has_null_key(self._h, ctypes.byref(n)))
        return bool(n.value)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            if self.ctx.alive:
                L.lib().gz_relation_release(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HashRelationKernel:
    @staticmethod
    def Make(ctx: Context, types: Sequence[int], key_col, builder_type: int = 1, normalize=None) -> HashRelation:
        return HashRelation(ctx, types, key_col, builder_type, normalize)


class ConditionedProbeKernel:
    def __init__(self, relation: HashRelation, probe_types: Sequence[int], probe_key: int, join_type: str,
                 null_aware: bool, result: Sequence[Tuple[int, int]], ctx: Optional[Context] = None):
        self.rel = relation
        self.ctx = ctx or relation.ctx
        self.probe_types = list(probe_types)
        self.result = [tuple(r) for r in result]
        self.result_types = [relation.types[i] if s == 0 else self.probe_types[i] for s, i in self.result]
        self._h = ctypes.c_void_p()
        sides = [s for s, _ in self.result]
        idx = [i for _, i in self.result]
        if isinstance(probe_key, (list, tuple)):
            L.check(L.lib().gz_probe_make_keys(relation.handle, len(probe_types), L.i32_array(probe_types),
                                               len(probe_key), L.i32_array(list(probe_key)), JOIN_TYPES[join_type],
                                               int(null_aware), len(self.result), L.i32_array(sides),
                                               L.i32_array(idx), ctypes.byref(self._h)))
        else:
            L.check(L.lib().gz_probe_make(relation.handle, len(probe_types), L.i32_array(probe_types), probe_key,
                                          JOIN_TYPES[join_type], int(null_aware), len(self.result),
                                          L.i32_array(sides), L.i32_array(idx), ctypes.byref(self._h)))
        if self.ctx is not relation.ctx:
            L.check(L.lib().gz_probe_set_context(self._h, self.ctx.handle))

    @classmethod
    def Make(cls, relation, probe_types, probe_key, join_type="inner", null_aware=False, result=(), ctx=None):
        """ctx: the probe's own context (stream and memory), for task threads that
        probe one shared relation at once; the relation's by default."""
        return cls(relation, probe_types, probe_key, join_type, null_aware, result, ctx)

    def Process(self, batch: DeviceBatch) -> int:
        self.ctx.hold(batch)
        n = ctypes.c_int64()
        L.check(L.lib().gz_probe_process(self._h, batch.num_rows, batch.addrs, batch.sizes, batch.num_bufs,
                                         ctypes.byref(n)))
        return n.value

    def ProcessBatches(self, batches: Sequence[DeviceBatch], arr=None) -> Tuple[int, np.ndarray]:
        self.ctx.hold(list(batches))
        if arr is None:
            arr = make_batch_array(batches)
        n = ctypes.c_int64()
        per = np.zeros(max(1, len(batches)), np.int64)
        L.check(L.lib().gz_probe_process_batches(self._h, len(batches), arr, ctypes.byref(n),
                                                 per.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return n.value, per[: len(batches)]

    def result_columns(self, n: int) -> List:
        """Host copy of the last output: per column (type, values, valid[, offsets, data]).
        The kernels run on the context's non-blocking stream, so the stream is
        drained before the null-stream copies read their output."""
        self.ctx.synchronize()
        out = []
        for c, t in enumerate(self.result_types):
            vals, valid, offs, data = (ctypes.c_void_p() for _ in range(4))
            L.check(L.lib().gz_probe_result_column(self._h, c, ctypes.byref(vals), ctypes.byref(valid),
                                                   ctypes.byref(offs), ctypes.byref(data)))
            ok = to_host(valid.value, n, np.uint8).astype(bool) if n else np.zeros(0, bool)
            if T.is_binary(t):
                o = to_host(offs.value, 4 * (n + 1), np.int32)
                d = to_host(data.value, int(o[-1])) if n and o[-1] > 0 else np.zeros(0, np.uint8)
                out.append((t, [bytes(d[o[i]:o[i + 1]]) for i in range(n)], ok))
            elif t == T.NA:
                out.append((t, [None] * n, ok))
            else:
                dt = np.dtype(_NP[t]) if t != T.DECIMAL128 else np.dtype("V16")
                out.append((t, to_host(vals.value, n * dt.itemsize, dt) if n else np.zeros(0, dt), ok))
        return out

    def close(self):
        if self._h:
            if self.ctx.alive and self.rel.ctx.alive:
                L.lib().gz_probe_release(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class JoinAggregateKernel:
    """The fused stage whole-stage codegen builds from conditionedProbeArraysInner
    feeding hashAggregateArrays (WholeStageCodeGenKernel,
    ext/whole_stage_codegen_kernel.cc:48-330): an inner probe of ``relation``
    whose joined rows update a hash aggregate directly, grouped by the join key.
    ``inputs``: the aggregate's input columns as join result columns
    [(side, index)] (side 0 build, 1 probe); ``actions`` index them, as the
    reference's aggregateExpressions do."""

    def __init__(self, relation: HashRelation, probe_types: Sequence[int], probe_key: int,
                 inputs: Sequence[Tuple[int, int]], actions, ctx=None, options=None):
        from .aggregate import AggOptions, HashAggregateKernel, HashAggregateResultIterator, _actions_c

        self.rel = relation
        self.probe_types = list(probe_types)
        self.inputs = [tuple(x) for x in inputs]
        types = [relation.types[i] if s == 0 else self.probe_types[i] for s, i in self.inputs]
        self.kernel = HashAggregateKernel(ctx or relation.ctx, types, actions, options or AggOptions())
        it = HashAggregateResultIterator.__new__(HashAggregateResultIterator)
        it.k = self.kernel
        it._h = ctypes.c_void_p()
        it.num_instances = 1
        it._offset = 0
        it._finished = False
        it._results = None
        opts = self.kernel.options.to_c()
        acts = _actions_c(self.kernel.actions)
        L.check(L.lib().gz_join_agg_make(relation.handle, (ctx or relation.ctx).handle, len(self.probe_types),
                                         L.i32_array(self.probe_types), probe_key, len(self.inputs),
                                         L.i32_array([s for s, _ in self.inputs]),
                                         L.i32_array([i for _, i in self.inputs]), len(self.kernel.actions), acts,
                                         ctypes.byref(opts), ctypes.byref(it._h)))
        self.iterator = it

    def ProcessBatches(self, batches: Sequence[DeviceBatch], arr=None):
        from .splitter import make_batch_array

        self.kernel.ctx.hold(list(batches))
        if arr is None:
            arr = make_batch_array(batches)
        L.check(L.lib().gz_join_agg_process_batches(self.iterator._h, len(batches), arr))

    def results(self):
        """(type, values, valid) per result column, groups in join-output order."""
        return self.iterator.results()

    def num_groups(self) -> int:
        return self.iterator.num_groups()

    def close(self):
        self.iterator.close()
