"""Result validity as Arrow bitmaps (what a JNI Next hands back in a
RecordBatch): the aggregate's, an aggregate set instance's, a multi-key
aggregate's and the probe's bitmaps equal the packed validity bytes, LSB
first, padded with 1 bits to a multiple of 64 rows; the generic packer does
the same into a caller buffer."""
import ctypes

import numpy as np
import pytest
import torch

from gazelle_plugin_amd import _lib as L
from gazelle_plugin_amd import datagen
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.aggregate import HashAggregateKernel, HashAggregateResultIterator
from gazelle_plugin_amd.batch import DeviceBatch, HostBatch, HostColumn, bitmap_from_bools
from gazelle_plugin_amd.hip import to_host
from gazelle_plugin_amd.join import ConditionedProbeKernel, HashRelationKernel
from aggcheck import CONFIG0_ACTIONS

pytestmark = pytest.mark.gpu


def _expect_bits(valid):
    return bitmap_from_bools(np.asarray(valid, bool))  # padded with 1s to 64-bit words


def _agg_batches(seed):
    rng = np.random.default_rng(seed)
    out = []
    for n in (3000, 777, 4096):
        key = rng.integers(0, 300, n).astype(np.int64)
        kv = rng.random(n) > 0.05  # a null key group
        price = rng.normal(size=n)
        pv = rng.random(n) > 0.4
        qty = rng.integers(1, 50, n).astype(np.int32)
        out.append(HostBatch(n, [HostColumn(T.INT64, key, bitmap_from_bools(kv)),
                                 HostColumn(T.DOUBLE, price, bitmap_from_bools(pv)), HostColumn(T.INT32, qty)]))
    return out


def test_aggregate_result_bitmaps(ctx):
    batches = _agg_batches(1)
    it = HashAggregateKernel.Make(ctx, datagen.AGG_TYPES, CONFIG0_ACTIONS).MakeResultIterator()
    it.ProcessBatches([DeviceBatch.from_host(b) for b in batches])
    res = it.results()
    n = len(res[0][1])
    for c, (_, _, valid) in enumerate(res):
        bm = ctypes.c_void_p()
        L.check(L.lib().gz_agg_result_bitmap(it._h, c, ctypes.byref(bm)))
        ctx.synchronize()
        got = to_host(bm.value, (n + 63) // 64 * 8)
        np.testing.assert_array_equal(got, _expect_bits(valid), err_msg=f"column {c}")
    assert not res[0][2].all()  # the null key group is there
    it.close()


def test_set_instance_and_multikey_bitmaps(ctx):
    inst = [_agg_batches(10 + i) for i in range(3)]
    dev, inst_of = [], []
    for i, bl in enumerate(inst):
        for b in bl:
            dev.append(DeviceBatch.from_host(b))
            inst_of.append(i)
    k = HashAggregateKernel.Make(ctx, datagen.AGG_TYPES, CONFIG0_ACTIONS)
    it = HashAggregateResultIterator(k, num_instances=len(inst))
    it.ProcessInstances(dev, inst_of)
    for i in range(len(inst)):
        res = it.results(i)
        n = len(res[0][1])
        for c, (_, _, valid) in enumerate(res):
            bm = ctypes.c_void_p()
            L.check(L.lib().gz_agg_set_result_bitmap(it._h, i, c, ctypes.byref(bm)))
            ctx.synchronize()
            np.testing.assert_array_equal(to_host(bm.value, (n + 63) // 64 * 8), _expect_bits(valid))
    it.close()
    mk = HashAggregateKernel.Make(ctx, datagen.AGG_TYPES, [("action_groupby", [0]), ("action_groupby", [2]),
                                                            ("action_sum", [1])]).MakeResultIterator()
    mk.ProcessBatches([DeviceBatch.from_host(b) for b in inst[0]])
    res = mk.results()
    n = len(res[0][1])
    for c, (_, _, valid) in enumerate(res):
        bm = ctypes.c_void_p()
        L.check(L.lib().gz_agg_result_bitmap(mk._h, c, ctypes.byref(bm)))
        ctx.synchronize()
        np.testing.assert_array_equal(to_host(bm.value, (n + 63) // 64 * 8), _expect_bits(valid))
    mk.close()


@pytest.mark.parametrize("join_type", ["inner", "semi"])
def test_probe_result_bitmaps(ctx, join_type):
    rng = np.random.default_rng(5)
    nb = 2000
    build = HostBatch(nb, [HostColumn(T.INT64, rng.permutation(5000)[:nb].astype(np.int64)),
                           HostColumn(T.DOUBLE, rng.normal(size=nb), bitmap_from_bools(rng.random(nb) > 0.3))])
    probe = HostBatch(7000, [HostColumn(T.INT64, rng.integers(0, 5000, 7000).astype(np.int64),
                                        bitmap_from_bools(rng.random(7000) > 0.02)),
                             HostColumn(T.INT32, rng.integers(0, 9, 7000).astype(np.int32),
                                        bitmap_from_bools(rng.random(7000) > 0.5))])
    rel = HashRelationKernel.Make(ctx, [T.INT64, T.DOUBLE], 0)
    rel.Evaluate(DeviceBatch.from_host(build))
    rel.Finish()
    result = [(1, 1), (0, 1)] if join_type == "inner" else [(1, 1), (1, 0)]
    pk = ConditionedProbeKernel.Make(rel, [T.INT64, T.INT32], 0, join_type, False, result)
    n = pk.Process(DeviceBatch.from_host(probe))
    cols = pk.result_columns(n)
    for c, (_, _, valid) in enumerate(cols):
        bm = ctypes.c_void_p()
        L.check(L.lib().gz_probe_result_bitmap(pk._h, c, ctypes.byref(bm)))
        ctx.synchronize()
        np.testing.assert_array_equal(to_host(bm.value, (n + 63) // 64 * 8), _expect_bits(valid))
    pk.close()
    rel.close()


def test_generic_validity_packer(ctx):
    for n in (1, 63, 64, 65, 1000):
        v = (np.random.default_rng(n).random(n) > 0.5).astype(np.uint8)
        dv = torch.from_numpy(v).to("cuda")
        out = torch.zeros((n + 63) // 64 * 8, dtype=torch.uint8, device="cuda")
        L.check(L.lib().gz_validity_bitmap(ctx.handle, ctypes.c_void_p(dv.data_ptr()), n,
                                           ctypes.c_void_p(out.data_ptr())))
        ctx.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), _expect_bits(v.astype(bool)))
