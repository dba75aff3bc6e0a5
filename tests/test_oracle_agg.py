"""Pins of the oracle's hash aggregate: the reference's own aggregate tests
(arrow_compute_test_aggregate.cc), brute force, and pyarrow's group_by."""
import math

import numpy as np
import pytest

import oracle
from gazelle_plugin_amd import datagen
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.aggregate import parse_action
from golden import fixtures as F
from aggcheck import CONFIG0_ACTIONS, python_groupby


def _same(a, b, tol):
    if b is None:
        return a is None
    if isinstance(b, str) and b == "NaN":
        return a is not None and math.isnan(a)
    if isinstance(b, str):
        return a == b
    if a is None or (isinstance(a, float) and math.isnan(a)):
        return False
    return abs(a - b) <= tol * max(1.0, abs(b))


def _column(res, i):
    t, v, ok = res[i]
    if t == T.STRING:
        return [v[j].decode() if ok[j] else None for j in range(len(v))]
    return [v[j].item() if ok[j] else None for j in range(len(v))]


@pytest.mark.parametrize("case", range(len(F.load("agg_fixtures.json")["tests"])))
def test_reference_aggregate_fixtures(case):
    fx = F.load("agg_fixtures.json")["tests"][case]
    schema = fx["schema"]
    batches = [F.batch(schema, cols) for cols in fx["batches"]]
    types = [F.type_id(t) for _, t in schema]
    actions = [parse_action(n, i) for n, i in fx["actions"]]
    res = oracle.aggregate(batches, types, actions)
    tol = fx.get("tolerance", 0.0)
    if "expected" in fx:
        for i, exp in enumerate(fx["expected"]):
            got = _column(res, i)
            assert len(got) == len(exp)
            assert all(_same(g, e, tol) for g, e in zip(got, exp)), (fx["name"], i, got, exp)
    else:
        e = fx["expected_named"]
        unique, s, cnt, mn, mx, avg = (_column(res, i) for i in range(6))
        assert unique == e["unique"]
        prod = [None if a is None else a * b for a, b in zip(s, cnt)]
        assert all(_same(g, x, 0) for g, x in zip(prod, e["sum_count_multiply"]))
        assert mn == e["min"]
        assert all(_same(g, x, tol) for g, x in zip(avg, e["avg"]))


def test_config0_query_brute_force():
    batches = datagen.agg_host(3000, num_keys=50, batch_rows=700)
    actions = [parse_action(n, i) for n, i in CONFIG0_ACTIONS]
    res = oracle.aggregate(batches, datagen.AGG_TYPES, actions)
    groups = python_groupby(batches, 0, [1, 2])
    keys = list(groups.keys())
    assert _column(res, 0) == keys  # first-seen order
    for j, k in enumerate(keys):
        price, qty = groups[k]
        pv = [p for p in price if p is not None]
        assert math.isclose(res[1][1][j], sum(pv), rel_tol=1e-12)
        assert math.isclose(res[2][1][j], sum(pv) / len(pv), rel_tol=1e-12)
        assert res[3][1][j] == len(price)
        assert res[4][1][j] == min(qty) and res[5][1][j] == max(qty)
        assert res[6][1][j] == sum(qty)


def test_config0_query_matches_pyarrow():
    pa = pytest.importorskip("pyarrow")
    import pyarrow.compute as pc
    from gazelle_plugin_amd.batch import to_arrow

    batches = datagen.agg_host(40000, num_keys=1000)
    actions = [parse_action(n, i) for n, i in CONFIG0_ACTIONS]
    res = oracle.aggregate(batches, datagen.AGG_TYPES, actions)
    tab = pa.Table.from_batches([to_arrow(b) for b in batches]).rename_columns(["k", "p", "q"])
    agg = tab.group_by("k", use_threads=False).aggregate(
        [("p", "sum"), ("p", "mean"), ("k", "count", pc.CountOptions("all")), ("q", "min"),
         ("q", "max"), ("q", "sum")]).sort_by("k")
    order = np.argsort(res[0][1], kind="stable")
    np.testing.assert_array_equal(res[0][1][order], agg["k"].to_numpy())
    np.testing.assert_allclose(res[1][1][order], agg["p_sum"].to_numpy(), rtol=1e-12)
    np.testing.assert_allclose(res[2][1][order], agg["p_mean"].to_numpy(), rtol=1e-12)
    np.testing.assert_array_equal(res[3][1][order], agg["k_count"].to_numpy())
    np.testing.assert_array_equal(res[4][1][order], agg["q_min"].to_numpy())
    np.testing.assert_array_equal(res[5][1][order], agg["q_max"].to_numpy())
    np.testing.assert_array_equal(res[6][1][order], agg["q_sum"].to_numpy())


def test_min_max_without_nan_check_follow_first_value():
    """action_min / action_max without the NaN check compare with < / >, so a
    leading NaN sticks and later NaNs are ignored (MinAction :1000-1016)."""
    from gazelle_plugin_amd.batch import HostBatch, HostColumn

    k = np.array([1, 1, 1, 2, 2, 2], np.int64)
    v = np.array([np.nan, 3.0, 1.0, 5.0, np.nan, 2.0])
    b = HostBatch(6, [HostColumn(T.INT64, k), HostColumn(T.DOUBLE, v)])
    res = oracle.aggregate([b], [T.INT64, T.DOUBLE],
                           [parse_action("action_groupby", [0]), parse_action("action_min", [1]),
                            parse_action("action_max", [1])])
    assert math.isnan(res[1][1][0]) and math.isnan(res[2][1][0])
    assert res[1][1][1] == 2.0 and res[2][1][1] == 5.0


MK_TYPES = [T.INT32, T.INT64, T.DOUBLE, T.DOUBLE, T.INT16]
MK_ACTIONS = [("action_groupby", [0]), ("action_sum", [2]), ("action_groupby", [1]),
              ("action_countLiteral_1", []), ("action_groupby", [3]), ("action_min", [4]),
              ("action_max", [4]), ("action_sum_count", [2]), ("action_count", [4])]


def test_multi_key_brute_force():
    """Several group-by keys: groups, their first-seen order, key values and
    aggregates against a brute force over key tuples (None for a null key, raw
    bits otherwise)."""
    from aggcheck import multikey_batches, python_groupby_keys
    batches = multikey_batches([700, 0, 1, 513], seed=3)
    res = oracle.aggregate(batches, MK_TYPES, [parse_action(n, i) for n, i in MK_ACTIONS])
    groups = python_groupby_keys(batches, [0, 1, 3], [2, 4])
    keys = list(groups)
    assert len(res) == 10 and len(res[0][1]) == len(keys)
    # key columns: result 0, 2, 4 (raw bits compared, nulls as None)
    for j, col in enumerate((0, 2, 4)):
        t, v, ok = res[col]
        bits = v.view(np.dtype(f"u{v.dtype.itemsize}"))
        assert [int(bits[i]) if ok[i] else None for i in range(len(v))] == [k[j] for k in keys]
    for i, k in enumerate(keys):
        vals, v4 = groups[k]
        nn = [x for x in vals if x is not None]
        nn4 = [x for x in v4 if x is not None]
        assert _same(_column(res, 1)[i], math.fsum(nn) if nn else None, 1e-12)
        assert _column(res, 3)[i] == len(vals)
        assert _column(res, 5)[i] == (min(nn4) if nn4 else None)
        assert _column(res, 6)[i] == (max(nn4) if nn4 else None)
        assert _same(_column(res, 7)[i], math.fsum(nn) if nn else None, 1e-12)
        assert _column(res, 8)[i] == (len(nn) if nn else None)
        assert _column(res, 9)[i] == len(nn4)


def test_multi_key_unsafe_row_equality_cases():
    """The UnsafeRow key compares the null bitset and the raw bytes of the
    non-null keys: the value under a null is ignored, (null, 5) and (5, null)
    differ although their appended bytes agree, -0.0 differs from 0.0, and NaNs
    differ by payload."""
    from gazelle_plugin_amd.batch import HostBatch, HostColumn, bitmap_from_bools
    k0 = np.array([5, 9, 5, 5, 7, 7, 7, 7, 7], np.int64)
    k0_ok = np.array([0, 0, 1, 1, 1, 1, 1, 1, 1], bool)
    k1 = np.array([5, 5, 3, 3, 0, 0, 0, 0, 0], np.int64)
    k1_ok = np.array([1, 1, 0, 0, 1, 1, 1, 1, 1], bool)
    d = np.array([0, 0, 0, 0, 0, 0x8000000000000000, 0, 0x7FF8000000000000, 0x7FF8000000000001],
                 np.uint64).view(np.float64)
    b = HostBatch(9, [HostColumn(T.INT64, k0, bitmap_from_bools(k0_ok)),
                      HostColumn(T.INT64, k1, bitmap_from_bools(k1_ok)),
                      HostColumn(T.DOUBLE, d, None)])
    acts = [("action_groupby", [0]), ("action_groupby", [1]), ("action_groupby", [2]),
            ("action_countLiteral_1", [])]
    res = oracle.aggregate([b], [T.INT64, T.INT64, T.DOUBLE], [parse_action(n, i) for n, i in acts])
    # rows 0,1: (null,5,0.0) twice (value under the null differs); rows 2,3:
    # (5,null,0.0) twice; row 4 and 6: (7,0,0.0); 5: -0.0; 7, 8: two NaNs
    assert _column(res, 3) == [2, 2, 2, 1, 1, 1]
    assert _column(res, 0) == [None, 5, 7, 7, 7, 7]
    assert _column(res, 1) == [5, None, 0, 0, 0, 0]
    bits = res[2][1].view(np.uint64).tolist()
    assert bits == [0, 0, 0, 0x8000000000000000, 0x7FF8000000000000, 0x7FF8000000000001]


def test_utf8_key_brute_force():
    """One utf8 key: groups by the bytes in first-seen order, the empty string
    and the null key each a group of their own (brute force over Python bytes);
    and (int32, utf8) keys against a brute force over (int or None, bytes or
    None) tuples."""
    from gazelle_plugin_amd.batch import HostBatch, HostColumn, bitmap_from_bools, bools_from_bitmap
    rng = np.random.default_rng(4)
    pool = [b"", b"a", b"ab", b"b", b"\x00", b"\x00\x00", b"abc" * 40]
    batches = []
    for n in (500, 0, 300):
        enc = [pool[i] for i in rng.integers(0, len(pool), n)]
        off = np.zeros(n + 1, np.int32)
        off[1:] = np.cumsum([len(e) for e in enc])
        data = np.frombuffer(b"".join(enc), np.uint8).copy() if off[-1] else np.zeros(0, np.uint8)
        batches.append(HostBatch(n, [HostColumn(T.INT32, rng.integers(0, 3, n).astype(np.int32),
                                                bitmap_from_bools(rng.random(n) > 0.2)),
                                     HostColumn(T.STRING, data, bitmap_from_bools(rng.random(n) > 0.2), off)]))
    rows = []
    for b in batches:
        kv = bools_from_bitmap(b.columns[0].validity, b.num_rows)
        sv = bools_from_bitmap(b.columns[1].validity, b.num_rows)
        o, d = b.columns[1].offsets, b.columns[1].values
        for r in range(b.num_rows):
            rows.append((int(b.columns[0].values[r]) if kv[r] else None,
                         d[o[r]:o[r + 1]].tobytes() if sv[r] else None))
    for keys, acts in (([1], [("action_groupby", [1]), ("action_countLiteral_1", [])]),
                       ([0, 1], [("action_groupby", [0]), ("action_groupby", [1]), ("action_countLiteral_1", [])])):
        groups = {}
        for row in rows:
            k = tuple(row[c] for c in keys)
            groups[k] = groups.get(k, 0) + 1
        res = oracle.aggregate(batches, [T.INT32, T.STRING], [parse_action(n, i) for n, i in acts])
        cols = [[res[j][1][i] if res[j][2][i] else None for i in range(len(res[-1][1]))] for j in range(len(keys))]
        got = list(zip(*cols))
        got = [tuple(x if not isinstance(x, np.integer) else int(x) for x in k) for k in got]
        assert got == list(groups), keys
        assert res[-1][1].tolist() == list(groups.values())


def test_unsafe_row_string_bytes_carry_no_length():
    """The reference appends a utf8 key's bytes to the UnsafeRow without their
    length (unsafe_row.h:126-137), so ("ab", "c") and ("a", "bc") are one group
    there; the oracle keeps that (the device rejects a utf8 key before another
    key, where this matters)."""
    from gazelle_plugin_amd.batch import HostBatch, HostColumn

    def col(strs):
        enc = [s.encode() for s in strs]
        off = np.zeros(len(enc) + 1, np.int32)
        off[1:] = np.cumsum([len(e) for e in enc])
        return HostColumn(T.STRING, np.frombuffer(b"".join(enc), np.uint8).copy(), None, off)

    b = HostBatch(3, [col(["ab", "a", "a"]), col(["c", "bc", "c"])])
    acts = [("action_groupby", [0]), ("action_groupby", [1]), ("action_countLiteral_1", [])]
    res = oracle.aggregate([b], [T.STRING, T.STRING], [parse_action(n, i) for n, i in acts])
    assert res[2][1].tolist() == [2, 1]
