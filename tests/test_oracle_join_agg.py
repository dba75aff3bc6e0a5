"""The expected results of the fused join -> aggregate (the oracle's aggregate
over the oracle's join output) pinned by the reference's whole-stage test
WSCGTestInnerJoinWithGroupbyAggregate (tests/golden/wscg_fixtures.json), and
by a brute-force nested-loop join + dictionary group-by on random inputs."""
import numpy as np
import pytest

from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.batch import HostBatch, HostColumn, bitmap_from_bools, bools_from_bitmap
from golden import fixtures as F
from joinaggcheck import expected, materialize


def test_reference_wscg_fixture():
    fx = F.load("wscg_fixtures.json")["tests"][0]
    build = [F.batch(fx["build_schema"], b) for b in fx["build"]]
    probe = [F.batch(fx["probe_schema"], b) for b in fx["probe"]]
    kt = F.type_id(fx["build_schema"][fx["build_key"]][1])
    inputs = [tuple(x) for x in fx["aggregate_inputs"]]
    # the join condition holds for every joined pair, so leaving it out changes nothing
    j = materialize(build, probe, 0, 0, kt, [(0, 1), (1, 1), (0, 2)])
    f1, t1f1, f2 = (np.asarray(c.values, np.int64) for c in j.columns)
    assert j.num_rows > 0 and bool(np.all(f1 + t1f1 > f2))
    got = expected(build, probe, 0, 0, kt, inputs, [tuple(a) for a in fx["actions"]])
    want_types = [F.type_id(t) for _, t in fx["result_schema"]]
    assert [t for t, _, _ in got] == want_types
    for (t, v, ok), e in zip(got, fx["expected"]):
        assert ok.all()
        assert v.tolist() == e


def test_random_against_brute_force():
    rng = np.random.default_rng(3)
    bkeys = rng.integers(0, 50, 300)
    bval = rng.integers(-100, 100, 300).astype(np.int32)
    bvalid = rng.random(300) > 0.1
    pkeys = rng.integers(0, 60, 2000)
    pvalid = rng.random(2000) > 0.05
    build = [HostBatch(300, [HostColumn(T.INT64, bkeys.astype(np.int64)),
                             HostColumn(T.INT32, bval, bitmap_from_bools(bvalid))])]
    probe = [HostBatch(1000, [HostColumn(T.INT64, pkeys[:1000].astype(np.int64), bitmap_from_bools(pvalid[:1000]))]),
             HostBatch(1000, [HostColumn(T.INT64, pkeys[1000:].astype(np.int64), bitmap_from_bools(pvalid[1000:]))])]
    actions = [("action_groupby", [0]), ("action_sum", [1]), ("action_count", [1]), ("action_countLiteral_1", []),
               ("action_min", [1]), ("action_max", [1])]
    got = expected(build, probe, 0, 0, T.INT64, [(1, 0), (0, 1)], actions)
    # brute force: nested loop join, then a dict group-by in first-seen order
    groups = {}
    for i in range(2000):
        if not pvalid[i]:
            continue
        for b in range(300):
            if bkeys[b] == pkeys[i]:
                g = groups.setdefault(int(pkeys[i]), [])
                g.append(int(bval[b]) if bvalid[b] else None)
    keys = list(groups)
    assert got[0][1].tolist() == keys
    vals = [groups[k] for k in keys]
    nn = [[x for x in v if x is not None] for v in vals]
    assert got[1][1][got[1][2]].tolist() == [sum(x) for x in nn if x]
    assert got[2][1].tolist() == [len(x) for x in nn]
    assert got[3][1].tolist() == [len(v) for v in vals]
    assert got[4][1][got[4][2]].tolist() == [min(x) for x in nn if x]
    assert got[5][1][got[5][2]].tolist() == [max(x) for x in nn if x]
