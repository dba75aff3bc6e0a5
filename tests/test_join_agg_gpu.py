"""Parity of the fused join -> aggregate (one pass: probe, then the aggregate
update of every joined pair, no join output) with the oracle's aggregate over
the oracle's join output: integer results bit-exact, FP64 within 1e-12
relative, groups in join-output order; the reference's whole-stage test
WSCGTestInnerJoinWithGroupbyAggregate reproduced."""
import numpy as np
import pytest

from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.batch import DeviceBatch, HostBatch, HostColumn, bitmap_from_bools
from gazelle_plugin_amd.join import HashRelationKernel, JoinAggregateKernel
from golden import fixtures as F
from aggcheck import assert_results_equal
from joinaggcheck import expected

pytestmark = pytest.mark.gpu

ACTIONS = [("action_groupby", [0]), ("action_sum", [1]), ("action_avg", [1]), ("action_countLiteral_1", []),
           ("action_min", [2]), ("action_max", [2]), ("action_sum", [2]), ("action_sum", [3]),
           ("action_count", [3]), ("action_sum_count", [1]), ("action_max_true", [1])]
INPUTS = [(1, 0), (0, 2), (0, 1), (1, 1)]  # probe key, build double, build int32, probe int32


def run(ctx, build, probe, btypes, ptypes, bkey, pkey, inputs, actions, calls=1, builder_type=1):
    rel = HashRelationKernel.Make(ctx, btypes, bkey, builder_type)
    rel.EvaluateBatches([DeviceBatch.from_host(b) for b in build])
    rel.Finish()
    k = JoinAggregateKernel(rel, ptypes, pkey, inputs, actions)
    dev = [DeviceBatch.from_host(b) for b in probe]
    per = (len(dev) + calls - 1) // calls
    for c in range(calls):
        part = dev[c * per:(c + 1) * per]
        if part:
            k.ProcessBatches(part)
    res = k.results()
    k.close()
    rel.close()
    return res


def make_build(rng, keys, null_frac=0.1):
    n = len(keys)
    dv = rng.standard_normal(n) * 100
    dv[rng.random(n) < 0.01] = np.nan
    return HostBatch(n, [HostColumn(T.INT64, keys.astype(np.int64)),
                         HostColumn(T.INT32, rng.integers(-1000, 1000, n).astype(np.int32),
                                    bitmap_from_bools(rng.random(n) > null_frac)),
                         HostColumn(T.DOUBLE, dv, bitmap_from_bools(rng.random(n) > null_frac))])


def make_probe(rng, keys, n, null_frac=0.02):
    return HostBatch(n, [HostColumn(T.INT64, keys.astype(np.int64), bitmap_from_bools(rng.random(n) > null_frac)),
                         HostColumn(T.INT32, rng.integers(-50, 50, n).astype(np.int32),
                                    bitmap_from_bools(rng.random(n) > 0.1))])


BT = [T.INT64, T.INT32, T.DOUBLE]
PT = [T.INT64, T.INT32]


def test_reference_wscg_fixture(ctx):
    fx = F.load("wscg_fixtures.json")["tests"][0]
    build = [F.batch(fx["build_schema"], b) for b in fx["build"]]
    probe = [F.batch(fx["probe_schema"], b) for b in fx["probe"]]
    btypes = [F.type_id(t) for _, t in fx["build_schema"]]
    ptypes = [F.type_id(t) for _, t in fx["probe_schema"]]
    got = run(ctx, build, probe, btypes, ptypes, 0, 0, [tuple(x) for x in fx["aggregate_inputs"]],
              [tuple(a) for a in fx["actions"]])
    assert [t for t, _, _ in got] == [F.type_id(t) for _, t in fx["result_schema"]]
    for (t, v, ok), e in zip(got, fx["expected"]):
        assert ok.all()
        assert v.tolist() == e


@pytest.mark.parametrize("calls", [1, 3])
def test_unique_build_keys(ctx, calls):
    rng = np.random.default_rng(5)
    bk = rng.permutation(np.arange(1, 40001)) * 7919
    build = [make_build(rng, bk[:25000]), make_build(rng, bk[25000:])]
    probe = []
    for n in (50000, 4096, 1, 70000):
        pk = np.where(rng.random(n) < 0.9, bk[(rng.zipf(1.3, n) - 1) % len(bk)], rng.integers(1, 10 ** 9, n))
        probe.append(make_probe(rng, pk, n))
    got = run(ctx, build, probe, BT, PT, 0, 0, INPUTS, ACTIONS, calls=calls)
    assert_results_equal(got, expected(build, probe, 0, 0, T.INT64, INPUTS, ACTIONS), 1e-12)


def test_repeated_build_keys_grouped_by_build_key(ctx):
    rng = np.random.default_rng(6)
    bk = rng.integers(0, 3000, 12000)
    build = [make_build(rng, bk)]
    probe = [make_probe(rng, rng.integers(0, 3500, 30000), 30000)]
    inputs = [(0, 0), (0, 2), (0, 1), (1, 1)]  # group by the build key column
    acts = [a for a in ACTIONS if a[0] != "action_max_true"]
    got = run(ctx, build, probe, BT, PT, 0, 0, inputs, acts)
    assert_results_equal(got, expected(build, probe, 0, 0, T.INT64, inputs, acts), 1e-12)


def test_hot_key_and_min_int64_key(ctx):
    rng = np.random.default_rng(7)
    bk = np.array([np.iinfo(np.int64).min, 5, 6, 7], np.int64)
    build = [make_build(rng, bk, null_frac=0.0)]
    pk = np.where(rng.random(60000) < 0.7, 5, np.where(rng.random(60000) < 0.5, np.iinfo(np.int64).min, 8))
    probe = [make_probe(rng, pk, 60000)]
    got = run(ctx, build, probe, BT, PT, 0, 0, INPUTS, ACTIONS)
    assert_results_equal(got, expected(build, probe, 0, 0, T.INT64, INPUTS, ACTIONS), 1e-12)


def test_high_cardinality_spills_past_lds(ctx):
    rng = np.random.default_rng(8)
    bk = np.arange(200000, dtype=np.int64) * 3 + 1
    build = [make_build(rng, bk)]
    probe = [make_probe(rng, rng.permutation(bk)[:150000], 150000), make_probe(rng, bk[rng.integers(0, 200000, 80000)], 80000)]
    got = run(ctx, build, probe, BT, PT, 0, 0, INPUTS, ACTIONS)
    assert_results_equal(got, expected(build, probe, 0, 0, T.INT64, INPUTS, ACTIONS), 1e-12)


def test_no_match_gives_no_groups(ctx):
    rng = np.random.default_rng(9)
    build = [make_build(rng, np.arange(100, dtype=np.int64))]
    probe = [make_probe(rng, np.arange(1000, 2000, dtype=np.int64), 1000)]
    got = run(ctx, build, probe, BT, PT, 0, 0, INPUTS, ACTIONS)
    assert all(len(v) == 0 for _, v, _ in got)
