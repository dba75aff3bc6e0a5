"""Parity of the HIP hash join with the oracle through the C ABI: output rows
bit-exact and in the reference's order (probe order, matches in build insertion
order)."""
import struct

import numpy as np
import pytest

import oracle
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.batch import DeviceBatch, HostBatch, HostColumn, bitmap_from_bools, bools_from_bitmap
from gazelle_plugin_amd.join import ConditionedProbeKernel, HashRelationKernel
from golden import fixtures as F
from test_oracle_join import JOIN_FIXTURES, JT, canon, fixture_keys, key_of, oracle_fixture_outputs, take

pytestmark = pytest.mark.gpu


def host_values(batch, c, rows):
    """Raw values + validity of rows (bytes for utf8)."""
    col = batch.columns[c]
    valid = bools_from_bitmap(col.validity, batch.num_rows)
    vals = []
    for r in rows:
        if col.type_id in (T.STRING, T.BINARY):
            vals.append(bytes(col.values[col.offsets[r]:col.offsets[r + 1]]))
        elif col.type_id == T.BOOL:
            vals.append(int(bools_from_bitmap(col.values, batch.num_rows)[r]))
        else:
            vals.append(col.values[r].item())
    return vals, [bool(valid[r]) for r in rows]


def check_output(got, exp_cols):
    """Column-by-column comparison that reports the first difference (a plain
    list assert makes pytest diff long lists for minutes)."""
    for c, ((t, vals, ok), (ev, eok)) in enumerate(zip(got, exp_cols)):
        ok = np.asarray(ok, bool)
        eok = np.asarray(eok, bool)
        if len(ok) != len(eok):
            raise AssertionError(f"column {c}: {len(ok)} rows, expected {len(eok)}")
        bad = np.nonzero(ok != eok)[0]
        if len(bad):
            raise AssertionError(f"column {c}: validity differs at {bad[:10].tolist()}")
        gv = [v.item() if hasattr(v, "item") else v for v in vals]
        for i in np.nonzero(eok)[0]:
            if isinstance(gv[i], float) and isinstance(ev[i], float):
                same = struct.pack("<d", gv[i]) == struct.pack("<d", ev[i])  # bit-exact, NaN included
            else:
                same = gv[i] == ev[i]
            if not same:
                raise AssertionError(f"column {c} row {i}: {gv[i]!r} != {ev[i]!r}")


def run_case(ctx, build, probe, btypes, ptypes, bkey, pkey, join_type, null_aware, result, batched,
             builder_type=1, normalize=None):
    rel = HashRelationKernel.Make(ctx, btypes, bkey, builder_type, normalize)
    if batched:
        rel.EvaluateBatches([DeviceBatch.from_host(b) for b in build])
    else:
        for b in build:
            rel.Evaluate(DeviceBatch.from_host(b))
    rel.Finish()
    pk = ConditionedProbeKernel.Make(rel, ptypes, pkey, join_type, null_aware, result)
    outs = []
    if batched:
        dev = [DeviceBatch.from_host(b) for b in probe]
        n, per = pk.ProcessBatches(dev)
        cols = pk.result_columns(n)
        off = 0
        for m in per:
            outs.append([(t, v[off:off + m], ok[off:off + m]) for t, v, ok in cols])
            off += m
    else:
        for b in probe:
            n = pk.Process(DeviceBatch.from_host(b))
            outs.append(pk.result_columns(n))
    pk.close()
    rel.close()
    return outs


def expected_cols(build, probe, bkey, pkey, ktype, join_type, null_aware, result, builder_type=1):
    bflat = []
    for bi, b in enumerate(build):
        bflat += [(bi, r) for r in range(b.num_rows)]
    outs = oracle.join([key_of(b, bkey) for b in build], [key_of(b, pkey) for b in probe], ktype,
                       JT[join_type], null_aware, builder_type)
    res = []
    for pb, out in zip(probe, outs):
        if join_type == "inner":
            prow, brow = out
        else:
            prow, brow = out, None
        cols = []
        for s, i in result:
            if s == 1:
                cols.append(host_values(pb, i, list(prow)))
            else:
                if brow is None:
                    cols.append(([None] * len(prow), [False] * len(prow)))
                else:
                    vals, oks = [], []
                    for g in brow:
                        bi, r = bflat[int(g)]
                        v, o = host_values(build[bi], i, [r])
                        vals += v
                        oks += o
                    cols.append((vals, oks))
        res.append(cols)
    return res


@pytest.mark.parametrize("case", range(len(JOIN_FIXTURES)))
@pytest.mark.parametrize("batched", [True, False])
def test_reference_join_fixtures_on_gpu(ctx, case, batched):
    """Every join golden, bit for bit (floats by their bits, NaN as NaN), and
    equal to the oracle's result on the same input."""
    fx = JOIN_FIXTURES[case]
    build = [F.batch(fx["build_schema"], b) for b in fx["build"]]
    probe = [F.batch(fx["probe_schema"], b) for b in fx["probe"]]
    btypes = [F.type_id(t) for _, t in fx["build_schema"]]
    ptypes = [F.type_id(t) for _, t in fx["probe_schema"]]
    bk, pk = fixture_keys(fx)
    outs = run_case(ctx, build, probe, btypes, ptypes, bk, pk, fx["join_type"], fx["null_aware"], fx["result"],
                    batched, normalize=fx.get("normalize"))
    assert len(outs) == len(fx["expected"])
    for got, exp, ora in zip(outs, fx["expected"], oracle_fixture_outputs(fx, build, probe)):
        vals = []
        for t, v, ok in got:
            vals.append([None if not o else (x.decode() if isinstance(x, bytes) else
                                             (bool(x) if t == T.BOOL else x.item()))
                         for x, o in zip(v, ok)])
        assert canon(vals) == canon(exp), (fx["name"], vals, exp)
        assert canon(vals) == canon(ora), (fx["name"], vals, ora)


@pytest.mark.parametrize("keys", ["one", "two"])
def test_normalized_float_keys_match_the_oracle(ctx, keys):
    """Spark's normalize on double keys: NaN, -0.0, tiny values and ordinary
    values, duplicates on the build side; one key (NaN never matches) and two
    keys (NaN matches NaN by bytes)."""
    rng = np.random.default_rng(17)
    special = np.array([np.nan, 0.0, -0.0, 1e-8, -5e-8, 1.0, -1.0, 2.5])
    def batch(n):
        k0 = special[rng.integers(0, len(special), n)]
        k1 = special[rng.integers(0, 3, n)]
        v = rng.random(n) > 0.1
        return HostBatch(n, [HostColumn(T.DOUBLE, k0.copy(), bitmap_from_bools(v)), HostColumn(T.DOUBLE, k1.copy()),
                             HostColumn(T.INT32, np.arange(n, dtype=np.int32))])
    build, probe = [batch(300), batch(77)], [batch(500), batch(1000)]
    types = [T.DOUBLE, T.DOUBLE, T.INT32]
    bk = 0 if keys == "one" else [0, 1]
    norm = [True] if keys == "one" else [True, True]
    for join_type in ("inner", "semi", "anti"):
        fx = {"build_key": bk, "probe_key": bk, "join_type": join_type, "null_aware": False, "normalize": norm,
              "build_schema": [["k0", "double"], ["k1", "double"], ["r", "int32"]],
              "result": [[1, 2], [0, 2]] if join_type == "inner" else [[1, 2]]}
        got = run_case(ctx, build, probe, types, types, bk, bk, join_type, False, fx["result"], True,
                       normalize=norm)
        exp = oracle_fixture_outputs(fx, build, probe)
        for g, e in zip(got, exp):
            vals = [[x.item() if o else None for x, o in zip(v, ok)] for _, v, ok in g]
            assert vals == e, (keys, join_type)


def _workload(rng, nbuild, nprobe, dup_keys, key_type=T.INT64, nulls=0.02):
    npt = {T.INT64: np.int64, T.INT32: np.int32, T.UINT16: np.uint16, T.DOUBLE: np.float64}[key_type]
    if dup_keys:
        bk = rng.integers(0, nbuild // 3 + 1, size=nbuild)
    else:
        bk = rng.permutation(nbuild * 4)[:nbuild]
    pk = np.where(rng.random(nprobe) < 0.9, bk[rng.integers(0, nbuild, size=nprobe)],
                  rng.integers(nbuild * 4, nbuild * 8, size=nprobe))
    bk, pk = bk.astype(npt), pk.astype(npt)
    if key_type == T.DOUBLE:
        pk[::97] = np.nan
    build, probe = [], []
    for b0 in range(0, nbuild, 3000):
        b1 = min(nbuild, b0 + 3000)
        n = b1 - b0
        v = rng.random(n) >= nulls
        build.append(HostBatch(n, [HostColumn(key_type, bk[b0:b1].copy(), bitmap_from_bools(v)),
                                   HostColumn(T.INT32, rng.integers(-1000, 1000, size=n).astype(np.int32)),
                                   HostColumn(T.DOUBLE, rng.standard_normal(n))]))
    for b0 in range(0, nprobe, 4096):
        b1 = min(nprobe, b0 + 4096)
        n = b1 - b0
        v = rng.random(n) >= nulls
        lens = rng.integers(0, 9, size=n).astype(np.int32)
        off = np.zeros(n + 1, np.int32)
        off[1:] = np.cumsum(lens)
        probe.append(HostBatch(n, [HostColumn(key_type, pk[b0:b1].copy(), bitmap_from_bools(v)),
                                   HostColumn(T.STRING, rng.integers(97, 123, size=int(off[-1]), dtype=np.uint8),
                                              None, off)]))
    return build, probe


@pytest.mark.parametrize("join_type", ["inner", "semi", "anti"])
@pytest.mark.parametrize("dup_keys", [False, True])
@pytest.mark.parametrize("batched", [True, False])
def test_random_joins_match_oracle(ctx, join_type, dup_keys, batched):
    rng = np.random.default_rng(hash((join_type, dup_keys)) % 1000)
    build, probe = _workload(rng, 10000, 20000, dup_keys)
    btypes, ptypes = [T.INT64, T.INT32, T.DOUBLE], [T.INT64, T.STRING]
    result = [[1, 0], [0, 1], [0, 2], [1, 1]] if join_type == "inner" else [[1, 0], [1, 1]]
    got = run_case(ctx, build, probe, btypes, ptypes, 0, 0, join_type, False, result, batched)
    exp = expected_cols(build, probe, 0, 0, T.INT64, join_type, False, result)
    for g, e in zip(got, exp):
        check_output(g, e)


@pytest.mark.parametrize("key_type", [T.INT32, T.UINT16, T.DOUBLE])
def test_key_types(ctx, key_type):
    rng = np.random.default_rng(key_type)
    build, probe = _workload(rng, 3000, 9000, True, key_type)
    btypes, ptypes = [key_type, T.INT32, T.DOUBLE], [key_type, T.STRING]
    for jt in ("inner", "semi", "anti"):
        result = [[1, 0], [0, 1]] if jt == "inner" else [[1, 0]]
        got = run_case(ctx, build, probe, btypes, ptypes, 0, 0, jt, False, result, True)
        exp = expected_cols(build, probe, 0, 0, key_type, jt, False, result)
        for g, e in zip(got, exp):
            check_output(g, e)


def test_null_aware_anti_and_semi_builder(ctx):
    rng = np.random.default_rng(9)
    build, probe = _workload(rng, 2000, 5000, True)
    btypes, ptypes = [T.INT64, T.INT32, T.DOUBLE], [T.INT64, T.STRING]
    # build side has null keys -> null-aware anti emits nothing
    got = run_case(ctx, build, probe, btypes, ptypes, 0, 0, "anti", True, [[1, 0]], True)
    assert all(len(g[0][1]) == 0 for g in got)
    # skip-duplicate builder (type 3) gives the same semi join
    got = run_case(ctx, build, probe, btypes, ptypes, 0, 0, "semi", False, [[1, 0], [1, 1]], True, 3)
    exp = expected_cols(build, probe, 0, 0, T.INT64, "semi", False, [[1, 0], [1, 1]], 3)
    for g, e in zip(got, exp):
        check_output(g, e)


def test_empty_build_side(ctx):
    build = [HostBatch(0, [HostColumn(T.INT64, np.zeros(0, np.int64)), HostColumn(T.INT32, np.zeros(0, np.int32)),
                           HostColumn(T.DOUBLE, np.zeros(0))])]
    rng = np.random.default_rng(1)
    _, probe = _workload(rng, 100, 1000, False)
    btypes, ptypes = [T.INT64, T.INT32, T.DOUBLE], [T.INT64, T.STRING]
    for jt, na, expect_all in (("inner", False, False), ("semi", False, False), ("anti", False, True),
                               ("anti", True, True)):
        got = run_case(ctx, build, probe, btypes, ptypes, 0, 0, jt, na, [[1, 0]], True)
        n = sum(len(g[0][1]) for g in got)
        assert n == (sum(b.num_rows for b in probe) if expect_all else 0)


@pytest.mark.parametrize("join_type,dup_keys,null_aware", [("inner", False, False), ("semi", False, False),
                                                           ("anti", False, False), ("anti", False, True),
                                                           ("semi", True, False), ("inner", True, False)])
def test_one_batch_calls_match_the_oracle(ctx, join_type, dup_keys, null_aware):
    """Per-batch calls with fixed-width outputs (a call of one batch of up to
    16384 rows runs as one launch of 256-row blocks that sum the emitted rows of
    the blocks before them): batch sizes around the block edges and past the
    largest such batch, checked row for row against the oracle."""
    rng = np.random.default_rng(7 + dup_keys)
    nbuild = 6000
    bk = rng.integers(0, nbuild // 3 + 1, size=nbuild) if dup_keys else rng.permutation(nbuild * 4)[:nbuild]
    build = []
    for b0 in range(0, nbuild, 2500):
        n = min(nbuild, b0 + 2500) - b0
        build.append(HostBatch(n, [HostColumn(T.INT64, bk[b0:b0 + n].astype(np.int64),
                                              bitmap_from_bools(rng.random(n) >= 0.01)),
                                   HostColumn(T.INT32, rng.integers(-1000, 1000, size=n).astype(np.int32),
                                              bitmap_from_bools(rng.random(n) >= 0.1)),
                                   HostColumn(T.DOUBLE, rng.standard_normal(n))]))
    probe = []
    for n in (1, 63, 64, 255, 256, 257, 1000, 4096, 16383, 16384, 16385, 0, 20000):
        pk = np.where(rng.random(n) < 0.8, bk[rng.integers(0, nbuild, size=n)], rng.integers(nbuild * 4, nbuild * 8, size=n))
        probe.append(HostBatch(n, [HostColumn(T.INT64, pk.astype(np.int64), bitmap_from_bools(rng.random(n) >= 0.02)),
                                   HostColumn(T.INT32, rng.integers(0, 1 << 30, size=n).astype(np.int32),
                                              bitmap_from_bools(rng.random(n) >= 0.05))]))
    btypes, ptypes = [T.INT64, T.INT32, T.DOUBLE], [T.INT64, T.INT32]
    result = [[1, 0], [0, 1], [0, 2], [1, 1]] if join_type == "inner" else [[1, 1], [1, 0]]
    got = run_case(ctx, build, probe, btypes, ptypes, 0, 0, join_type, null_aware, result, False)
    exp = expected_cols(build, probe, 0, 0, T.INT64, join_type, null_aware, result)
    assert len(got) == len(exp)
    for g, e in zip(got, exp):
        check_output(g, e)


def test_task_threads_probe_one_relation_on_their_own_contexts(ctx):
    """A relation shared by several task threads (a broadcast relation): each
    thread's probe runs on its own context, per batch, at the same time as the
    others; every thread's output equals the oracle's for its batches."""
    import threading

    from gazelle_plugin_amd.context import Context

    rng = np.random.default_rng(11)
    build, probe = _workload(rng, 8000, 4 * 8 * 4096, False)
    probe = [HostBatch(b.num_rows, [b.columns[0]]) for b in probe]  # the key only: fixed-width outputs
    btypes = [T.INT64, T.INT32, T.DOUBLE]
    rel = HashRelationKernel.Make(ctx, btypes, 0)
    rel.EvaluateBatches([DeviceBatch.from_host(b) for b in build])
    rel.Finish()
    result = [[1, 0], [0, 1], [0, 2]]
    exp = expected_cols(build, probe, 0, 0, T.INT64, "inner", False, result)
    nthreads = 4
    ctxs = [Context(0) for _ in range(nthreads)]
    pks = [ConditionedProbeKernel.Make(rel, [T.INT64], 0, "inner", False, result, ctx=c) for c in ctxs]
    dev = [DeviceBatch.from_host(b) for b in probe]
    got = [None] * len(probe)
    errors = []

    def task(t):
        try:
            for i in range(t, len(probe), nthreads):
                n = pks[t].Process(dev[i])
                got[i] = pks[t].result_columns(n)
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=task, args=(t,)) for t in range(nthreads)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    for g, e in zip(got, exp):
        check_output(g, e)
    for pk in pks:
        pk.close()
    for c in ctxs:
        c.close()
    rel.close()


def test_probe_context_is_set_before_the_first_call(ctx):
    from gazelle_plugin_amd import _lib as L
    from gazelle_plugin_amd.context import Context

    b = HostBatch(3, [HostColumn(T.INT64, np.array([1, 2, 3], np.int64))])
    rel = HashRelationKernel.Make(ctx, [T.INT64], 0)
    rel.Evaluate(DeviceBatch.from_host(b))
    rel.Finish()
    pk = ConditionedProbeKernel.Make(rel, [T.INT64], 0, "semi", False, [[1, 0]])
    assert pk.Process(DeviceBatch.from_host(b)) == 3
    other = Context(0)
    assert L.lib().gz_probe_set_context(pk._h, other.handle) == 4  # Invalid
    pk.close()
    other.close()
    rel.close()



@pytest.mark.parametrize("join_type", ["inner", "semi", "anti"])
def test_batched_probe_many_tiles(ctx, join_type):
    """Batched probes whose rows emit at most one row each: 100-200 tiles of
    4096 rows per call (every tile's output offset from the scan of the tile
    counts), with null probe keys, three calls on one probe handle.  Expected
    rows from the oracle's join, columns gathered with numpy."""
    rng = np.random.default_rng({"inner": 71, "semi": 72, "anti": 73}[join_type])
    bk = rng.permutation(1 << 20)[:50000].astype(np.int64)
    bint = rng.integers(-1000, 1000, size=len(bk)).astype(np.int32)
    bdbl = rng.standard_normal(len(bk))
    cuts = ((0, 20000), (20000, 30000))
    build = [HostBatch(n, [HostColumn(T.INT64, bk[b0:b0 + n].copy()), HostColumn(T.INT32, bint[b0:b0 + n].copy()),
                           HostColumn(T.DOUBLE, bdbl[b0:b0 + n].copy())]) for b0, n in cuts]
    btypes, ptypes = [T.INT64, T.INT32, T.DOUBLE], [T.INT64, T.INT32]
    result = [[1, 0], [0, 1], [0, 2], [1, 1]] if join_type == "inner" else [[1, 0], [1, 1]]
    rel = HashRelationKernel.Make(ctx, btypes, 0)
    for b in build:
        rel.Evaluate(DeviceBatch.from_host(b))
    rel.Finish()
    pr = ConditionedProbeKernel.Make(rel, ptypes, 0, join_type, False, result)
    for call in range(3):
        probe = []
        for _ in range(int(rng.integers(40, 70))):
            n = int(rng.integers(9000, 13000))
            pk = np.where(rng.random(n) < 0.7, bk[rng.integers(0, len(bk), n)], rng.integers(0, 1 << 20, n))
            probe.append(HostBatch(n, [HostColumn(T.INT64, pk.astype(np.int64), bitmap_from_bools(rng.random(n) >= 0.01)),
                                       HostColumn(T.INT32, rng.integers(0, 1 << 30, size=n).astype(np.int32))]))
        n, per = pr.ProcessBatches([DeviceBatch.from_host(b) for b in probe])
        got = pr.result_columns(n)
        outs = oracle.join([key_of(b, 0) for b in build], [key_of(b, 0) for b in probe], T.INT64,
                           JT[join_type], False, 1)
        ek, ekv, ei, ebi, ebd = [], [], [], [], []
        for pb, out in zip(probe, outs):
            prow = np.asarray(out[0] if join_type == "inner" else out, np.int64)
            ek.append(pb.columns[0].values[prow])
            ekv.append(bools_from_bitmap(pb.columns[0].validity, pb.num_rows)[prow])
            ei.append(pb.columns[1].values[prow])
            if join_type == "inner":
                brow = np.asarray(out[1], np.int64)
                ebi.append(bint[brow])
                ebd.append(bdbl[brow])
        assert [int(x) for x in per] == [len(x) for x in ek]
        exp_cols = [(np.concatenate(ek), np.concatenate(ekv))]
        if join_type == "inner":
            exp_cols += [(np.concatenate(ebi), None), (np.concatenate(ebd), None)]
        exp_cols.append((np.concatenate(ei), None))
        assert n == len(exp_cols[0][0])
        for (t, vals, ok), (ev, eok) in zip(got, exp_cols):
            ok = np.asarray(ok, bool)
            if eok is None:
                assert ok.all()
            else:
                np.testing.assert_array_equal(ok, eok)
            vals = np.asarray(vals)
            m = ok if eok is not None else np.ones(len(ok), bool)
            np.testing.assert_array_equal(vals[m].view(np.uint8), np.asarray(ev)[m].view(np.uint8))
    pr.close()
    rel.close()
