"""Parity of the HIP hash aggregate with the oracle through the C ABI: integer
results bit-exact, FP64 sums/averages within 1e-12 relative, groups in the
reference's first-seen order."""
import numpy as np
import pytest

import oracle
from gazelle_plugin_amd import datagen
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.aggregate import AggOptions, HashAggregateKernel, parse_action
from gazelle_plugin_amd.batch import DeviceBatch, HostBatch, HostColumn
from golden import fixtures as F
from aggcheck import CONFIG0_ACTIONS, assert_results_equal
from splitcheck import random_batches

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def run_gpu(ctx, batches, types, actions, batched=True, **opts):
    k = HashAggregateKernel.Make(ctx, types, actions, AggOptions(**opts))
    it = k.MakeResultIterator()
    dev = [DeviceBatch.from_host(b) for b in batches]
    if batched:
        it.ProcessBatches(dev)
    else:
        for d in dev:
            it.ProcessAndCacheOne(d)
    res = it.results()
    it.close()
    return res


def expected(batches, types, actions):
    return oracle.aggregate(batches, types, [parse_action(n, i) for n, i in actions])


@pytest.mark.parametrize("batched", [True, False])
@pytest.mark.parametrize("placement", [0, 1, 2, 3])
def test_config0_query_small(ctx, batched, placement):
    batches = datagen.agg_host(50000, num_keys=10000)
    got = run_gpu(ctx, batches, datagen.AGG_TYPES, CONFIG0_ACTIONS, batched, placement=placement)
    assert_results_equal(got, expected(batches, datagen.AGG_TYPES, CONFIG0_ACTIONS), RTOL)


@pytest.mark.parametrize("num_keys", [1, 7, 200, 3000, 60000])
def test_cardinalities(ctx, num_keys):
    batches = datagen.agg_host(70000, num_keys=num_keys, batch_rows=9000)
    got = run_gpu(ctx, batches, datagen.AGG_TYPES, CONFIG0_ACTIONS)
    assert_results_equal(got, expected(batches, datagen.AGG_TYPES, CONFIG0_ACTIONS), RTOL)


@pytest.mark.parametrize("case", range(len(F.load("agg_fixtures.json")["tests"])))
def test_reference_fixtures_on_gpu(ctx, case):
    fx = F.load("agg_fixtures.json")["tests"][case]
    schema = fx["schema"]
    batches = [F.batch(schema, cols) for cols in fx["batches"]]
    types = [F.type_id(t) for _, t in schema]
    for batched in (True, False):
        got = run_gpu(ctx, batches, types, fx["actions"], batched)
        assert_results_equal(got, expected(batches, types, fx["actions"]), RTOL)


def test_nulls_key_types_and_all_actions(ctx):
    types = [T.INT32, T.DOUBLE, T.INT64, T.UINT64, T.FLOAT, T.INT16]
    batches = random_batches(types, [3000, 1, 0, 2500], seed=5, null_frac=0.2)
    for b in batches:  # a small key domain with nulls
        if b.num_rows:
            b.columns[0].values = (b.columns[0].values % 37).astype(np.int32)
            b.columns[2].values = (b.columns[2].values % 100000)
            b.columns[1].values[::7] = np.nan
    actions = [("action_groupby", [0]), ("action_sum", [1]), ("action_avg", [1]), ("action_count", [1]),
               ("action_countLiteral_1", []), ("action_min_true", [1]), ("action_max_true", [1]),
               ("action_min", [1]), ("action_max", [1]), ("action_sum", [2]), ("action_min", [3]),
               ("action_max", [3]), ("action_sum_count", [4]), ("action_min", [5]), ("action_max", [2]),
               ("action_avgByCount", [1, 2])]
    exp = expected(batches, types, actions)
    for placement in (1, 2, 3):
        got = run_gpu(ctx, batches, types, actions, placement=placement)
        assert_results_equal(got, exp, RTOL)


def test_global_aggregate(ctx):
    batches = datagen.agg_host(20000, num_keys=10)
    actions = [("action_sum", [1]), ("action_avg", [1]), ("action_countLiteral_1", []),
               ("action_min", [2]), ("action_max", [2])]
    got = run_gpu(ctx, batches, datagen.AGG_TYPES, actions)
    assert_results_equal(got, expected(batches, datagen.AGG_TYPES, actions), RTOL)


def test_next_batches_of_ten_thousand(ctx):
    batches = datagen.agg_host(60000, num_keys=25000)
    k = HashAggregateKernel.Make(ctx, datagen.AGG_TYPES, CONFIG0_ACTIONS)
    it = k.MakeResultIterator()
    it.ProcessBatches([DeviceBatch.from_host(b) for b in batches])
    sizes = []
    while it.HasNext():
        sizes.append(len(it.Next()[0][1]))
    it.close()
    n = len(expected(batches, datagen.AGG_TYPES, CONFIG0_ACTIONS)[0][1])
    assert sum(sizes) == n and all(s == 10000 for s in sizes[:-1])


def test_aggregate_set_instances_independent(ctx):
    """Many independent GROUP BYs in one launch equal each run alone."""
    from gazelle_plugin_amd.aggregate import HashAggregateResultIterator

    rng = np.random.default_rng(3)
    inst_batches = []
    for i in range(12):
        rows = int(rng.integers(1000, 30000))
        keys = int(rng.integers(5, 20000))
        inst_batches.append(datagen.agg_host(rows, num_keys=keys, batch_rows=4096, seed=100 + i))
    k = HashAggregateKernel.Make(ctx, datagen.AGG_TYPES, CONFIG0_ACTIONS)
    it = HashAggregateResultIterator(k, num_instances=len(inst_batches))
    dev, inst_of = [], []
    for i, bl in enumerate(inst_batches):
        for b in bl:
            dev.append(DeviceBatch.from_host(b))
            inst_of.append(i)
    it.ProcessInstances(dev, inst_of)
    for i, bl in enumerate(inst_batches):
        assert_results_equal(it.results(i), expected(bl, datagen.AGG_TYPES, CONFIG0_ACTIONS), RTOL)
    it.close()


def _instances(seed, count=10):
    rng = np.random.default_rng(seed)
    inst_batches = []
    for i in range(count):
        rows = int(rng.integers(500, 30000))
        keys = int(rng.integers(3, 20000))
        inst_batches.append(datagen.agg_host(rows, num_keys=keys, batch_rows=4096, seed=seed * 100 + i))
    dev, inst_of = [], []
    for i, bl in enumerate(inst_batches):
        for b in bl:
            dev.append(DeviceBatch.from_host(b))
            inst_of.append(i)
    return inst_batches, dev, inst_of


def test_planned_set_run_matches_oracle(ctx):
    """prepare + run (kernels only) gives each instance's oracle result, twice in a row."""
    from gazelle_plugin_amd.aggregate import HashAggregateResultIterator

    inst_batches, dev, inst_of = _instances(5)
    k = HashAggregateKernel.Make(ctx, datagen.AGG_TYPES, CONFIG0_ACTIONS)
    it = HashAggregateResultIterator(k, num_instances=len(inst_batches))
    it.Prepare(dev, inst_of)
    exp = [expected(bl, datagen.AGG_TYPES, CONFIG0_ACTIONS) for bl in inst_batches]
    for _ in range(2):
        it.Run()
        for i in range(len(inst_batches)):
            assert_results_equal(it.results(i), exp[i], RTOL)
    it.close()


def test_graph_replays_match_eager(ctx):
    """The captured graph replayed several times, with eager work and host
    synchronization between replays, gives the eager run's results every time."""
    import torch

    from gazelle_plugin_amd.aggregate import HashAggregateResultIterator

    inst_batches, dev, inst_of = _instances(6)
    k = HashAggregateKernel.Make(ctx, datagen.AGG_TYPES, CONFIG0_ACTIONS)
    it = HashAggregateResultIterator(k, num_instances=len(inst_batches))
    it.Prepare(dev, inst_of)
    it.Run()
    eager = [it.results(i) for i in range(len(inst_batches))]
    it.Capture()
    for rep in range(3):
        x = torch.ones(1 << 20, device="cuda") * rep  # unrelated eager work between replays
        it.Launch()
        ctx.synchronize()
        torch.cuda.synchronize()
        for i in range(len(inst_batches)):
            assert_results_equal(it.results(i), eager[i], RTOL)
        del x
    it.close()


def test_autotuned_plan_matches_oracle(ctx):
    from gazelle_plugin_amd.aggregate import HashAggregateResultIterator

    inst_batches, dev, inst_of = _instances(7, count=6)
    k = HashAggregateKernel.Make(ctx, datagen.AGG_TYPES, CONFIG0_ACTIONS)
    it = HashAggregateResultIterator(k, num_instances=len(inst_batches))
    chosen = it.Autotune(dev, inst_of)
    assert chosen >= 0
    it.Capture()
    it.Launch()
    for i, bl in enumerate(inst_batches):
        assert_results_equal(it.results(i), expected(bl, datagen.AGG_TYPES, CONFIG0_ACTIONS), RTOL)
    it.close()


@pytest.mark.parametrize("batched", [True, False])
def test_high_cardinality_requested_partitioned(ctx, batched):
    """300k groups on the partitioned placement: 1024 partitions of ~300
    groups, past the small-fan-out scatter; the results equal the oracle's."""
    batches = datagen.agg_host(330000, num_keys=300000, batch_rows=8192)
    got = run_gpu(ctx, batches, datagen.AGG_TYPES, CONFIG0_ACTIONS, batched, placement=3)
    assert_results_equal(got, expected(batches, datagen.AGG_TYPES, CONFIG0_ACTIONS), RTOL)


def _fmix64(x):
    """The table hash of the device aggregate (MurmurHash3's 64-bit finalizer)."""
    x = x.astype(np.uint64)
    x ^= x >> np.uint64(33)
    x *= np.uint64(0xFF51AFD7ED558CCD)
    x ^= x >> np.uint64(33)
    x *= np.uint64(0xC4CEB9FE1A85EC53)
    x ^= x >> np.uint64(33)
    return x


@pytest.mark.parametrize("batched", [True, False])
def test_owned_partition_whose_lds_table_overflows(ctx, batched):
    """Two partitions (the high hash bit) of 900-1023 groups each in 1024-slot
    tables, ~12k records each: each partition is one work item, whose LDS
    table (linear probing, at most 32 probes) overflows at that load; the item
    then resets its partition's HBM slots, stores its LDS groups by claiming
    slots and applies the records of the keys it lacks with atomics.  The
    results equal the oracle's."""
    actions = [("action_groupby", [0]), ("action_sum", [1]), ("action_countLiteral_1", [])]
    batches = datagen.agg_host(24000, num_keys=1900, seed=9)
    keys = np.unique(np.concatenate([b.columns[0].values for b in batches]))
    counts = np.bincount((_fmix64(keys.view(np.uint64)) >> np.uint64(63)).astype(np.int64), minlength=2)
    assert counts.min() >= 900 and counts.max() < 1024, counts
    got = run_gpu(ctx, batches, datagen.AGG_TYPES, actions, batched, placement=3, expected_groups=1000)
    assert_results_equal(got, expected(batches, datagen.AGG_TYPES, actions), RTOL)


@pytest.mark.parametrize("placement", [0, 2, 3])
def test_table_overflow_is_reported(ctx, placement):
    """A cardinality hint (500) far below the true 20000 groups overflows the
    HBM table sized from it: the call or the finish reports CapacityError."""
    from gazelle_plugin_amd._lib import GazelleError

    bl = datagen.agg_host(60000, num_keys=20000, seed=25)
    k = HashAggregateKernel.Make(ctx, datagen.AGG_TYPES, CONFIG0_ACTIONS,
                                 AggOptions(placement=placement, expected_groups=500))
    it = k.MakeResultIterator()
    with pytest.raises(GazelleError, match="overflow"):
        it.ProcessBatches([DeviceBatch.from_host(b) for b in bl])
        it.finish()
    it.close()
