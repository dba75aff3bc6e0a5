"""The batched aggregate at the batched workload's shape, against the oracle.

64 independent GROUP BYs (one RecordBatch each) with row counts log-uniform over
1k-1M and cardinalities log-uniform over 10-10^7 (capped at the rows), plus two
instances of 1M rows with 600k and 150k groups, so that the placements run
inside one set: the LDS first level and the partitioned placement, whose
fan-out reaches 2048 partitions of 512 groups (the HBM table takes only
instances past 4096 such partitions).  The set runs on the default
plan (every placement), then autotuned, captured as a HIP graph and replayed
(the tuner may pick any placements); every instance equals the
oracle's aggregate of its batch (integers and min/max bit for bit, FP64 sums
and averages within 1e-12)."""
import numpy as np
import pytest

import oracle
from gazelle_plugin_amd import datagen
from gazelle_plugin_amd.aggregate import HashAggregateKernel, HashAggregateResultIterator, parse_action
from gazelle_plugin_amd.batch import DeviceBatch, HostBatch, HostColumn, bitmap_from_bools
from gazelle_plugin_amd import types as T
from aggcheck import CONFIG0_ACTIONS, assert_results_equal

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def instance_batch(rows, groups, seed):
    """The batched workload's columns: int64 key over `groups` values, float64
    price ~ N(100, 25) with 1% nulls, int32 qty ~ U[1, 50]."""
    rng = np.random.default_rng(seed)
    keyset = rng.integers(-(1 << 62), 1 << 62, size=groups, dtype=np.int64)
    key = keyset[rng.integers(0, groups, size=rows)]
    price = rng.normal(100.0, 25.0, size=rows)
    valid = rng.random(rows) >= 0.01
    qty = rng.integers(1, 51, size=rows).astype(np.int32)
    return HostBatch(rows, [HostColumn(T.INT64, key), HostColumn(T.DOUBLE, price, bitmap_from_bools(valid)),
                            HostColumn(T.INT32, qty)])


def test_batched_workload_shape_matches_oracle(ctx):
    rows, groups = datagen.batched_agg_shapes(64, seed=7)
    rows = list(rows) + [1_000_000, 1_000_000]
    groups = list(groups) + [600_000, 150_000]
    host = [instance_batch(int(r), int(g), 1000 + i) for i, (r, g) in enumerate(zip(rows, groups))]
    dev = [DeviceBatch.from_host(b) for b in host]
    k = HashAggregateKernel.Make(ctx, datagen.AGG_TYPES, CONFIG0_ACTIONS)
    acts = [parse_action(n, i) for n, i in CONFIG0_ACTIONS]
    exp = [oracle.aggregate([b], datagen.AGG_TYPES, acts) for b in host]
    # the default plan places instances by their estimated groups: the LDS first
    # level and the partitioned placement (small and large fan-outs) in one set
    it = HashAggregateResultIterator(k, num_instances=len(host))
    it.Prepare(dev, list(range(len(host))))
    placements = {it.placement(i) for i in range(len(host))}
    assert placements == {1, 3}, placements
    it.Run()
    for i in range(len(host)):
        assert_results_equal(it.results(i), exp[i], RTOL)
    it.close()
    # the autotuned plan (whichever placements won the timing), captured and replayed
    it = HashAggregateResultIterator(k, num_instances=len(host))
    it.Autotune(dev, list(range(len(host))))
    it.Capture()
    for rep in range(2):
        it.Launch()
        ctx.synchronize()
        for i in range(len(host)):
            got = it.results(i)
            assert len(got[0][1]) == len(exp[i][0][1]), (rep, i, rows[i], groups[i])
            assert_results_equal(got, exp[i], RTOL)
    it.close()
