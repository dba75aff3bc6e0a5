"""Floating min / max and float group keys on the GPU, bit for bit against the
oracle: the sign of a zero and the payload of a NaN come out as the reference
keeps them (first zero seen; first or last NaN by the action's rule), in every
table placement (LDS first level, HBM, partitioned), in the per-batch and
batched entry points, with and without a group-by key, in aggregate sets (run
and graph replay) and in the fused join -> aggregate."""
import numpy as np
import pytest

import oracle
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.aggregate import AggOptions, HashAggregateKernel, HashAggregateResultIterator, parse_action
from gazelle_plugin_amd.batch import DeviceBatch, HostBatch, HostColumn, bitmap_from_bools
from aggcheck import assert_results_equal

pytestmark = pytest.mark.gpu

D_POOL = np.array([0x0, 0x8000000000000000, 0x3FF0000000000000, 0xBFF0000000000000, 0x4000000000000000,
                   0x7FF8000000000000, 0x7FF8000000000ABC, 0xFFF8000000000001, 0x7FF0000000000123],
                  np.uint64).view(np.float64)
F_POOL = np.array([0x0, 0x80000000, 0x3F800000, 0xBF800000, 0x40000000, 0x7FC00000, 0x7FC00ABC, 0xFFC00001,
                   0x7FC00123], np.uint32).view(np.float32)

MINMAX = [("action_min", [1]), ("action_max", [1]), ("action_min_true", [1]), ("action_max_true", [1]),
          ("action_min", [2]), ("action_max", [2]), ("action_min_true", [2]), ("action_max_true", [2]),
          ("action_sum", [1]), ("action_count", [2])]
TYPES = [T.INT64, T.DOUBLE, T.FLOAT]


def batches(seed, sizes, groups, null_frac=0.15, special=0.5):
    """key int64 over `groups` values; a double and a float column whose values
    are +-0, +-1, 2 and NaN payloads (quiet and signalling) with probability
    `special`, else normal."""
    rng = np.random.default_rng(seed)
    out = []
    for n in sizes:
        key = rng.integers(0, groups, n).astype(np.int64) * 7 - 3
        sp = rng.random(n) < special
        d = np.where(sp, D_POOL[rng.integers(0, len(D_POOL), n)], rng.standard_normal(n))
        f = np.where(sp, F_POOL[rng.integers(0, len(F_POOL), n)], rng.standard_normal(n).astype(np.float32))
        out.append(HostBatch(n, [HostColumn(T.INT64, key),
                                 HostColumn(T.DOUBLE, d.astype(np.float64), bitmap_from_bools(rng.random(n) >= null_frac)),
                                 HostColumn(T.FLOAT, f.astype(np.float32), bitmap_from_bools(rng.random(n) >= null_frac))]))
    return out


def run(ctx, bl, types, actions, batched=True, **opts):
    it = HashAggregateKernel.Make(ctx, types, actions, AggOptions(**opts)).MakeResultIterator()
    dev = [DeviceBatch.from_host(b) for b in bl]
    if batched:
        it.ProcessBatches(dev)
    else:
        for d in dev:
            it.ProcessAndCacheOne(d)
    res = it.results()
    it.close()
    return res


def expected(bl, types, actions):
    return oracle.aggregate(bl, types, [parse_action(n, i) for n, i in actions])


@pytest.mark.parametrize("placement", [0, 1, 2, 3])
@pytest.mark.parametrize("batched", [True, False])
def test_grouped_min_max_bits(ctx, placement, batched):
    bl = batches(1, [9000, 1, 20000, 4096], 150)
    acts = [("action_groupby", [0])] + MINMAX
    assert_results_equal(run(ctx, bl, TYPES, acts, batched, placement=placement), expected(bl, TYPES, acts))


@pytest.mark.parametrize("placement", [0, 2, 3])
def test_grouped_min_max_bits_many_groups(ctx, placement):
    """~30k groups: the partitioned placement's fan-out and the HBM table."""
    bl = batches(2, [70000, 50000], 30000, special=0.7)
    acts = [("action_groupby", [0])] + MINMAX
    assert_results_equal(run(ctx, bl, TYPES, acts, placement=placement), expected(bl, TYPES, acts))


@pytest.mark.parametrize("batched", [True, False])
def test_global_min_max_bits(ctx, batched):
    """No group-by key: the reference's Evaluate(in) rules (min always NaN-safe;
    max with NaN check keeps the first NaN of the last batch holding one)."""
    for seed in range(4):
        bl = batches(10 + seed, [300, 50, 700, 2], 1, special=0.3)
        assert_results_equal(run(ctx, bl, TYPES, MINMAX, batched), expected(bl, TYPES, MINMAX))


@pytest.mark.parametrize("batched", [True, False])
def test_global_integer_min_max_all_null_batch(ctx, batched):
    """An all-null batch after a batch with values merges 0 (the null scalar's
    value) into the min / max; an all-null batch first does nothing."""
    def b(vals, nulls=False):
        n = len(vals)
        return HostBatch(n, [HostColumn(T.INT32, np.array(vals, np.int32),
                                        bitmap_from_bools(np.zeros(n, bool)) if nulls else None),
                             HostColumn(T.UINT64, np.array(vals, np.int64).astype(np.uint64),
                                        bitmap_from_bools(np.zeros(n, bool)) if nulls else None)])
    acts = [("action_min", [0]), ("action_max", [0]), ("action_min", [1]), ("action_max", [1]),
            ("action_countLiteral_1", [])]
    ty = [T.INT32, T.UINT64]
    for bl in ([b([5, 7]), b([1, 1, 3], True)], [b([1, 1], True), b([5, 7])], [b([-5, -7]), b([1], True)],
               [b([5, 9]), b([4, 4], True), b([2, 8])]):
        assert_results_equal(run(ctx, bl, ty, acts, batched), expected(bl, ty, acts))


@pytest.mark.parametrize("ktype", [T.DOUBLE, T.FLOAT])
@pytest.mark.parametrize("placement", [0, 1, 2, 3])
def test_float_group_key(ctx, ktype, placement):
    """-0.0 and 0.0 one group, every NaN one group, the key as its first row's bits."""
    rng = np.random.default_rng(20)
    pool = D_POOL if ktype == T.DOUBLE else F_POOL
    bl = []
    for n in (5000, 3, 12000):
        sp = rng.random(n) < 0.6
        other = rng.integers(0, 500, n).astype(np.float64) / 4
        k = np.where(sp, pool[rng.integers(0, len(pool), n)], other).astype(pool.dtype)
        bl.append(HostBatch(n, [HostColumn(ktype, k, bitmap_from_bools(rng.random(n) >= 0.05)),
                                HostColumn(T.INT64, rng.integers(-100, 100, n).astype(np.int64)),
                                HostColumn(T.DOUBLE, D_POOL[rng.integers(0, len(D_POOL), n)])]))
    ty = [ktype, T.INT64, T.DOUBLE]
    acts = [("action_groupby", [0]), ("action_sum", [1]), ("action_count", [1]), ("action_min", [2]),
            ("action_max_true", [2]), ("action_min", [1])]
    for batched in (True, False):
        assert_results_equal(run(ctx, bl, ty, acts, batched, placement=placement), expected(bl, ty, acts))


def test_set_run_and_graph_replay_bits(ctx):
    rng = np.random.default_rng(30)
    inst = [batches(40 + i, [int(rng.integers(100, 20000))], int(rng.integers(1, 3000))) for i in range(8)]
    dev, inst_of = [], []
    for i, bl in enumerate(inst):
        for b in bl:
            dev.append(DeviceBatch.from_host(b))
            inst_of.append(i)
    acts = [("action_groupby", [0])] + MINMAX
    k = HashAggregateKernel.Make(ctx, TYPES, acts)
    it = HashAggregateResultIterator(k, num_instances=len(inst))
    it.Prepare(dev, inst_of)
    exp = [expected(bl, TYPES, acts) for bl in inst]
    it.Run()
    for i in range(len(inst)):
        assert_results_equal(it.results(i), exp[i])
    it.Capture()
    for _ in range(3):
        it.Launch()
        ctx.synchronize()
        for i in range(len(inst)):
            assert_results_equal(it.results(i), exp[i])
    it.close()


def test_fused_join_aggregate_bits(ctx):
    """Float min / max over a probe column (repeated build keys) and a build
    column (unique build keys) of the fused join -> aggregate."""
    from gazelle_plugin_amd.join import HashRelationKernel, JoinAggregateKernel
    from joinaggcheck import expected as jexpected

    rng = np.random.default_rng(50)
    for unique in (True, False):
        nb = 3000
        bk = rng.permutation(nb).astype(np.int64) if unique else rng.integers(0, 800, nb).astype(np.int64)
        build = [HostBatch(nb, [HostColumn(T.INT64, bk),
                                HostColumn(T.DOUBLE, D_POOL[rng.integers(0, len(D_POOL), nb)],
                                           bitmap_from_bools(rng.random(nb) > 0.1))])]
        probe = []
        for n in (20000, 777):
            probe.append(HostBatch(n, [HostColumn(T.INT64, rng.integers(0, 1000, n).astype(np.int64)),
                                       HostColumn(T.DOUBLE, D_POOL[rng.integers(0, len(D_POOL), n)],
                                                  bitmap_from_bools(rng.random(n) > 0.1))]))
        if unique:
            inputs = [(1, 0), (0, 1), (1, 1)]
            acts = [("action_groupby", [0]), ("action_min", [1]), ("action_max_true", [1]),
                    ("action_min_true", [2]), ("action_max", [2]), ("action_countLiteral_1", [])]
        else:
            inputs = [(1, 0), (1, 1)]
            acts = [("action_groupby", [0]), ("action_min", [1]), ("action_max", [1]), ("action_min_true", [1]),
                    ("action_max_true", [1]), ("action_countLiteral_1", [])]
        rel = HashRelationKernel.Make(ctx, [T.INT64, T.DOUBLE], 0, 1)
        rel.EvaluateBatches([DeviceBatch.from_host(b) for b in build])
        rel.Finish()
        k = JoinAggregateKernel(rel, [T.INT64, T.DOUBLE], 0, inputs, acts)
        for b in probe:
            k.ProcessBatches([DeviceBatch.from_host(b)])
        got = k.results()
        k.close()
        rel.close()
        assert_results_equal(got, jexpected(build, probe, 0, 0, T.INT64, inputs, acts))
