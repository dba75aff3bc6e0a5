"""Python mirror of the reference hash aggregate over the C ABI.

Mirrors ``HashAggregateKernel::Make`` / ``MakeResultIterator`` and
``HashAggregateResultIterator<number>`` (ext/hash_aggregate_kernel.cc:147-224,
:778-928): actions are named as the reference names them (``action_groupby``,
``action_sum``, ``action_avg``, ``action_countLiteral_1``, ``action_min_true``,
``action_max``, ``action_sum_count``, ``action_avgByCount``, ``action_count``);
``ProcessAndCacheOne`` consumes one batch, ``HasNext`` / ``Next`` emit the groups
in first-seen order, ``NATIVESQL_BATCH_SIZE`` (10000) groups per batch.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib as L
from . import types as T
from .batch import DeviceBatch
from .context import Context
from .hip import to_host
from .splitter import make_batch_array

GROUPBY, COUNT, COUNT_LITERAL, SUM, AVG, MIN, MAX, SUM_COUNT, AVG_BY_COUNT = range(9)

_NP = {T.UINT8: np.uint8, T.INT8: np.int8, T.UINT16: np.uint16, T.INT16: np.int16,
       T.UINT32: np.uint32, T.INT32: np.int32, T.UINT64: np.uint64, T.INT64: np.int64,
       T.FLOAT: np.float32, T.DOUBLE: np.float64, T.DATE32: np.int32, T.DATE64: np.int64,
       T.TIMESTAMP: np.int64, T.BOOL: np.uint8}


def parse_action(name: str, inputs: Sequence[int]) -> Tuple[int, int, int, int]:
    """Reference action name -> (kind, input0, input1, nan_check)
    (PrepareActionList, hash_aggregate_kernel.cc:673-773)."""
    i0 = inputs[0] if len(inputs) > 0 else -1
    i1 = inputs[1] if len(inputs) > 1 else -1
    if name == "action_groupby":
        return (GROUPBY, i0, -1, 0)
    if name == "action_count":
        return (COUNT, i0, i1, 0)
    if name.startswith("action_countLiteral_"):
        return (COUNT_LITERAL, -1, -1, 0)
    if name == "action_sum":
        return (SUM, i0, -1, 0)
    if name == "action_avg":
        return (AVG, i0, -1, 0)
    if name == "action_sum_count":
        return (SUM_COUNT, i0, -1, 0)
    if name == "action_avgByCount":
        return (AVG_BY_COUNT, i0, i1, 0)
    for prefix, kind in (("action_min", MIN), ("action_max", MAX)):
        if name == prefix or name.startswith(prefix + "_"):
            nan = name[len(prefix) + 1:] == "true" if name != prefix else False
            return (kind, i0, -1, int(nan))
    raise NotImplementedError(f"{name} is not implemented.")


def batch_size() -> int:
    """NATIVESQL_BATCH_SIZE (ext/codegen_common.cc:556-565), default 10000."""
    try:
        return int(os.environ.get("NATIVESQL_BATCH_SIZE", "10000"))
    except ValueError:
        return 10000


@dataclass
class AggOptions:
    placement: int = 0        # 0 chosen per instance, 1 LDS first level, 2 HBM only, 3 partitioned
    tile_rows: int = 0
    load_factor: float = 0.5
    expected_groups: int = 0

    def to_c(self):
        o = L.GzAggOptions()
        o.placement = self.placement
        o.tile_rows = self.tile_rows
        o.load_factor = self.load_factor
        o.expected_groups = self.expected_groups
        return o


def _actions_c(actions):
    arr = (L.GzAggAction * len(actions))()
    for i, (k, a, b, n) in enumerate(actions):
        arr[i].kind, arr[i].input0, arr[i].input1, arr[i].nan_check = k, a, b, n
    return arr


class HashAggregateKernel:
    def __init__(self, ctx: Context, input_types: Sequence[int], actions, options: Optional[AggOptions] = None):
        self.ctx = ctx
        self.input_types = list(input_types)
        self.actions = [parse_action(n, i) if isinstance(n, str) else (n, *i) for n, i in actions]
        self.options = options or AggOptions()

    @classmethod
    def Make(cls, ctx: Context, input_types: Sequence[int], actions, options=None) -> "HashAggregateKernel":
        """actions: [(reference action name, [input column indices]), ...]."""
        return cls(ctx, input_types, actions, options)

    def MakeResultIterator(self) -> "HashAggregateResultIterator":
        return HashAggregateResultIterator(self)


class HashAggregateResultIterator:
    def __init__(self, kernel: HashAggregateKernel, num_instances: int = 1):
        self.k = kernel
        self._h = ctypes.c_void_p()
        self.num_instances = num_instances
        acts = _actions_c(kernel.actions)
        opts = kernel.options.to_c()
        types = L.i32_array(kernel.input_types)
        if num_instances == 1:
            L.check(L.lib().gz_agg_make(kernel.ctx.handle, len(kernel.input_types), types, len(kernel.actions),
                                        acts, ctypes.byref(opts), ctypes.byref(self._h)))
        else:
            L.check(L.lib().gz_agg_set_make(kernel.ctx.handle, len(kernel.input_types), types, len(kernel.actions),
                                            acts, ctypes.byref(opts), num_instances, ctypes.byref(self._h)))
        self._offset = 0
        self._finished = False
        self._results = None

    # -- input
    def ProcessAndCacheOne(self, batch: DeviceBatch):
        self.k.ctx.hold(batch)
        L.check(L.lib().gz_agg_process(self._h, batch.num_rows, batch.addrs, batch.sizes, batch.num_bufs))

    def ProcessBatches(self, batches: Sequence[DeviceBatch], arr=None):
        self.k.ctx.hold(list(batches))
        if arr is None:
            arr = make_batch_array(batches)
        L.check(L.lib().gz_agg_process_batches(self._h, len(batches), arr))

    def ProcessInstances(self, batches: Sequence[DeviceBatch], instance_of: Sequence[int], arr=None):
        self.k.ctx.hold(list(batches))
        if arr is None:
            arr = make_batch_array(batches)
        L.check(L.lib().gz_agg_set_process(self._h, len(batches), arr, L.i32_array(instance_of)))

    # -- planned form: prepare once, then run / capture + launch (HIP graph)
    def Prepare(self, batches: Sequence[DeviceBatch], instance_of: Sequence[int], arr=None):
        self._plan_batches = list(batches)  # a plan reads its batches at every run / launch
        self._plan_arr = arr if arr is not None else make_batch_array(batches)  # must outlive the plan
        self._plan_inst = L.i32_array(instance_of)
        L.check(L.lib().gz_agg_set_prepare(self._h, len(batches), self._plan_arr, self._plan_inst))
        self._finished = False

    def Autotune(self, batches: Sequence[DeviceBatch], instance_of: Sequence[int], arr=None) -> int:
        self._plan_batches = list(batches)
        self._plan_arr = arr if arr is not None else make_batch_array(batches)
        self._plan_inst = L.i32_array(instance_of)
        chosen = ctypes.c_int32()
        L.check(L.lib().gz_agg_set_autotune(self._h, len(batches), self._plan_arr, self._plan_inst,
                                            ctypes.byref(chosen)))
        self._finished = False
        return chosen.value

    def Run(self):
        L.check(L.lib().gz_agg_set_run(self._h))
        self._finished = False
        self._results = None

    def Capture(self):
        L.check(L.lib().gz_agg_set_capture(self._h))

    def Launch(self):
        L.check(L.lib().gz_agg_set_launch(self._h))
        self._finished = False
        self._results = None

    # -- output
    def finish(self):
        if not self._finished:
            L.check(L.lib().gz_agg_finish(self._h))
            self._finished = True

    def num_groups(self, inst: int = 0) -> int:
        self.finish()
        n = ctypes.c_int64()
        if self.num_instances == 1:
            L.check(L.lib().gz_agg_num_groups(self._h, ctypes.byref(n)))
        else:
            L.check(L.lib().gz_agg_set_num_groups(self._h, inst, ctypes.byref(n)))
        return n.value

    def placement(self, inst: int = 0) -> int:
        out = ctypes.c_int32()
        L.check(L.lib().gz_agg_set_placement(self._h, inst, ctypes.byref(out)))
        return out.value

    def results(self, inst: int = 0) -> List[Tuple[int, np.ndarray, np.ndarray]]:
        """All result columns of an instance: (type, values, valid) per column."""
        self.finish()
        self.k.ctx.synchronize()
        g = self.num_groups(inst)
        ncol = L.lib().gz_agg_num_result_columns(self._h)
        out = []
        for c in range(ncol):
            t = L.lib().gz_agg_result_type(self._h, c)
            if t == T.STRING:  # a utf8 key: bytes objects per group (see valid)
                offp, datap, validp = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
                L.check(L.lib().gz_agg_result_binary(self._h, c, ctypes.byref(offp), ctypes.byref(datap),
                                                     ctypes.byref(validp)))
                offs = to_host(offp.value, (g + 1) * 4, np.dtype(np.int32))
                nbytes = int(offs[-1])
                data = to_host(datap.value, nbytes, np.dtype(np.uint8)) if nbytes else np.zeros(0, np.uint8)
                ok = to_host(validp.value, g, np.uint8).astype(bool) if g else np.zeros(0, bool)
                v = np.array([data[offs[i]:offs[i + 1]].tobytes() for i in range(g)] + [None], object)[:g]
                out.append((t, v, ok))
                continue
            vals, valid = ctypes.c_void_p(), ctypes.c_void_p()
            if self.num_instances == 1:
                L.check(L.lib().gz_agg_result_column(self._h, c, ctypes.byref(vals), ctypes.byref(valid)))
            else:
                L.check(L.lib().gz_agg_set_result_column(self._h, inst, c, ctypes.byref(vals), ctypes.byref(valid)))
            dt = np.dtype(_NP[t])
            v = to_host(vals.value, g * dt.itemsize, dt) if g else np.zeros(0, dt)
            ok = to_host(valid.value, g, np.uint8).astype(bool) if g else np.zeros(0, bool)
            out.append((t, v, ok))
        return out

    def HasNext(self) -> bool:
        self.finish()
        return self._offset < self.num_groups()

    def Next(self):
        """Next NATIVESQL_BATCH_SIZE groups: [(type, values, valid)] per column."""
        if self._results is None:
            self._results = self.results()
        bs = batch_size()
        a, b = self._offset, min(self._offset + bs, self.num_groups())
        self._offset = b
        return [(t, v[a:b], ok[a:b]) for t, v, ok in self._results]

    def close(self):
        if self._h:
            if self.k.ctx.alive:
                L.lib().gz_agg_release(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
