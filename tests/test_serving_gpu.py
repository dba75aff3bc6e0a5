"""The fragment server: every fragment's result is the same when it runs alone
and when it runs in the mix (concurrent slots, Poisson arrivals); aggregate
results also match the oracle.  Tolerance: integers bit-exact, FP64 sums and
averages within 1e-12 relative (atomic summation order differs between runs)."""
import numpy as np
import pytest

import oracle
from gazelle_plugin_amd import datagen
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.aggregate import parse_action
from gazelle_plugin_amd.batch import DeviceBatch, HostBatch, HostColumn
from gazelle_plugin_amd.join import HashRelationKernel
from gazelle_plugin_amd.serving import AGGREGATE, JOIN_AGGREGATE, SPLIT, FragmentServer, latency_summary
from gazelle_plugin_amd.splitter import make_batch_array
from aggcheck import CONFIG0_ACTIONS, assert_results_equal
from joinaggcheck import expected as joinagg_expected
from splitcheck import assert_layout_equal

pytestmark = pytest.mark.gpu
RTOL = 1e-12
JOIN_RESULT = [(1, 0), (0, 2), (0, 1)]  # probe fk int64, build f64, build int32: the aggregate's schema


def _inputs(ctx):
    rng = np.random.default_rng(11)
    agg_host = datagen.agg_host(60 * 1024, num_keys=3000, batch_rows=1024)
    split_host = datagen.split_host(40 * 1024, batch_rows=1024)
    nb = 20000
    bk = rng.permutation(1 << 20)[:nb].astype(np.int64)
    build = [HostBatch(nb, [HostColumn(T.INT64, bk), HostColumn(T.INT32, rng.integers(-99, 99, nb).astype(np.int32)),
                            HostColumn(T.DOUBLE, rng.standard_normal(nb))])]
    probe_host = []
    for b in range(40):
        fk = np.where(rng.random(1024) < 0.95, bk[rng.integers(0, nb, 1024)], rng.integers(1 << 21, 1 << 22, 1024))
        probe_host.append(HostBatch(1024, [HostColumn(T.INT64, fk.astype(np.int64))]))
    rel = HashRelationKernel.Make(ctx, [T.INT64, T.INT32, T.DOUBLE], 0)
    rel.EvaluateBatches([DeviceBatch.from_host(b) for b in build])
    rel.Finish()
    dev = {k: [DeviceBatch.from_host(b) for b in v] for k, v in
           (("agg", agg_host), ("split", split_host), ("probe", probe_host))}
    return agg_host, split_host, probe_host, build, rel, dev


def _server(slots):
    return FragmentServer(0, 8 << 30, slots, agg=(datagen.AGG_TYPES, CONFIG0_ACTIONS),
                          split=(datagen.SPLIT_TYPES, 16, [0]),
                          probe=([T.INT64], 0, "inner", JOIN_RESULT),
                          join_agg=([T.INT64, T.DOUBLE, T.INT32], CONFIG0_ACTIONS))


def test_results_alone_equal_results_in_the_mix(ctx):
    agg_host, split_host, probe_host, build, rel, dev = _inputs(ctx)
    arrays = {k: make_batch_array(v) for k, v in dev.items()}
    rng = np.random.default_rng(12)
    frags, t = [], 0.0
    kinds = [AGGREGATE, JOIN_AGGREGATE, SPLIT]
    total = {AGGREGATE: len(agg_host), JOIN_AGGREGATE: len(probe_host), SPLIT: len(split_host)}
    name = {AGGREGATE: "agg", JOIN_AGGREGATE: "probe", SPLIT: "split"}
    for i in range(24):
        k = kinds[i % 3]
        nb = int(rng.integers(1, 12))
        b0 = int(rng.integers(0, total[k] - nb + 1))
        t += float(rng.exponential(0.0005))
        frags.append((k, 0, arrays[name[k]], b0, nb, t))
    srv = _server(4)
    rid = srv.add_relation(rel)
    assert rid == 0
    stats = srv.run(frags, keep=True)
    assert np.all(stats["status"] == 0)
    assert np.all(stats["admitted_s"] >= stats["arrival_s"] - 1e-9)
    assert np.all(stats["done_s"] >= stats["admitted_s"])
    assert len(set(stats["slot"].tolist())) > 1  # the mix really used several slots
    mixed = []
    for i, (k, _, _, b0, nb, _) in enumerate(frags):
        mixed.append(srv.split_result(i) if k == SPLIT else srv.agg_result(i, None, 0))
    alone_srv = _server(1)
    alone_srv.add_relation(rel)
    for i, f in enumerate(frags):
        k, _, _, b0, nb, _ = f
        st = alone_srv.run([(k, 0, f[2], b0, nb, 0.0)], keep=True)
        assert st["status"][0] == 0
        if k == SPLIT:
            assert_layout_equal(alone_srv.split_result(0), mixed[i], datagen.SPLIT_TYPES)
            exp = oracle.split(split_host[b0:b0 + nb], datagen.SPLIT_TYPES, 16, [0], "hash", 42)
            assert_layout_equal(mixed[i], exp, datagen.SPLIT_TYPES)
        else:
            assert_results_equal(mixed[i], alone_srv.agg_result(0, None, 0), RTOL)
            if k == AGGREGATE:
                exp = oracle.aggregate(agg_host[b0:b0 + nb], datagen.AGG_TYPES,
                                       [parse_action(n, j) for n, j in CONFIG0_ACTIONS])
                assert_results_equal(mixed[i], exp, RTOL)
            else:  # join then aggregate: the oracle's aggregate over the oracle's join output
                exp = joinagg_expected(build, probe_host[b0:b0 + nb], 0, 0, T.INT64, JOIN_RESULT, CONFIG0_ACTIONS)
                assert_results_equal(mixed[i], exp, RTOL)
    summ = latency_summary(stats)
    assert summ["p99_ms"] >= summ["p50_ms"] > 0
    alone_srv.close()
    srv.close()
    rel.close()


def test_admission_respects_the_budget(ctx):
    """With a budget that fits one fragment's estimate at a time, fragments
    never overlap even with several idle slots."""
    agg_host, split_host, probe_host, build, rel, dev = _inputs(ctx)
    arr = make_batch_array(dev["agg"])
    srv = FragmentServer(0, 100 << 20, 4, agg=(datagen.AGG_TYPES, CONFIG0_ACTIONS))
    frags = [(AGGREGATE, 0, arr, 4 * i, 4, 0.0) for i in range(6)]
    stats = srv.run(frags)
    assert np.all(stats["status"] == 0)
    order = np.argsort(stats["admitted_s"])
    for a, b in zip(order[:-1], order[1:]):
        assert stats["admitted_s"][b] >= stats["done_s"][a] - 1e-6
    srv.close()
    rel.close()
