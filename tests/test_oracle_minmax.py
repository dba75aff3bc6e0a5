"""Pins of the oracle's floating min/max and float group keys, bit for bit.

The reference keeps one input value per group (ext/actions_impl.cc MinAction
:905-936 / MaxAction :1504-1539 for grouped rows; GetMinResult :976-993 /
GetMaxResult :1581-1621 for the aggregate without a group-by key), so ties
between -0.0 and 0.0 and NaN payloads are decided by input order.  The checks
below state the result in closed form, derived once from those loops, instead
of replaying them:

grouped, NaN_check (min):   first non-NaN value equal to the least non-NaN value;
                            else the first NaN.
grouped, NaN_check (max):   the LAST NaN if any (every NaN overwrites);
                            else the first value equal to the greatest.
grouped, plain compares:    the first value if it is NaN (a NaN state never
                            compares); else the first non-NaN value equal to the
                            least (greatest) non-NaN value.
no group-by key, min:       as grouped NaN_check min, whatever NaN_check says.
no group-by key, max+check: the first NaN of the LAST batch holding a NaN;
                            else the first value equal to the greatest.
no group-by key, integers:  arrow MinMax per batch; an all-null batch after a
                            batch with values merges its null scalar's value 0.

A float group key (SparseHashMap<double>, sparse_hash_map.h:86-160; phmap
Hash<double>, phmap_utils.h:251-258): -0.0 and 0.0 are one group, every NaN is
one group, and the key comes out as the group's first row's bits.
"""
import math
import struct

import numpy as np
import pytest

import oracle
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.aggregate import parse_action
from gazelle_plugin_amd.batch import HostBatch, HostColumn, bitmap_from_bools

POS0, NEG0 = 0x0000000000000000, 0x8000000000000000
NANS = [0x7FF8000000000000, 0x7FF8000000000ABC, 0xFFF8000000000001, 0x7FF0000000000123]  # incl. a signalling NaN
POOL = [POS0, NEG0, 0x3FF0000000000000, 0xBFF0000000000000] + NANS  # +-0, +-1, four NaNs


def _f(bits):
    return struct.unpack("<d", struct.pack("<Q", bits))[0]


def _isnan(bits):
    return (bits & 0x7FF0000000000000) == 0x7FF0000000000000 and (bits & 0xFFFFFFFFFFFFF) != 0


def _first_equal(vals, target):
    return next(b for b in vals if not _isnan(b) and _f(b) == target)


def expect_grouped(vals, kind, nan_check):
    """vals: the group's non-null input bits in row order."""
    if not vals:
        return None
    non_nan = [b for b in vals if not _isnan(b)]
    nan = [b for b in vals if _isnan(b)]
    if kind == "min":
        if not nan_check and _isnan(vals[0]):
            return vals[0]
        return _first_equal(vals, min(_f(b) for b in non_nan)) if non_nan else nan[0]
    if nan_check:
        return nan[-1] if nan else _first_equal(vals, max(_f(b) for b in non_nan))
    if _isnan(vals[0]):
        return vals[0]
    return _first_equal(vals, max(_f(b) for b in non_nan))


def expect_global(batches_vals, kind, nan_check):
    """batches_vals: per batch, the non-null input bits in row order."""
    vals = [b for bv in batches_vals for b in bv]
    if not vals:
        return None
    if kind == "min":
        return expect_grouped(vals, "min", True)
    if nan_check:
        with_nan = [bv for bv in batches_vals if any(_isnan(b) for b in bv)]
        if with_nan:
            return next(b for b in with_nan[-1] if _isnan(b))
    return expect_grouped(vals, "max", False)


def _batches(rng, nb, n, ngroups, null_frac=0.2):
    out = []
    for _ in range(nb):
        key = rng.integers(0, ngroups, n).astype(np.int64)
        bits = np.array([POOL[i] for i in rng.integers(0, len(POOL), n)], np.uint64)
        ok = rng.random(n) >= null_frac
        out.append(HostBatch(n, [HostColumn(T.INT64, key), HostColumn(T.DOUBLE, bits.view(np.float64),
                                                                      bitmap_from_bools(ok))]))
    return out, ok


def _values(b):
    from gazelle_plugin_amd.batch import bools_from_bitmap
    ok = bools_from_bitmap(b.columns[1].validity, b.num_rows)
    return b.columns[0].values, b.columns[1].values.view(np.uint64), ok


@pytest.mark.parametrize("seed", range(6))
def test_grouped_float_min_max_bits(seed):
    rng = np.random.default_rng(seed)
    batches, _ = _batches(rng, 3, 40, 6)
    acts = [("action_groupby", [0]), ("action_min", [1]), ("action_max", [1]),
            ("action_min_true", [1]), ("action_max_true", [1])]
    res = oracle.aggregate(batches, [T.INT64, T.DOUBLE], [parse_action(n, i) for n, i in acts])
    groups = {}
    for b in batches:
        keys, bits, ok = _values(b)
        for r in range(b.num_rows):
            g = groups.setdefault(int(keys[r]), [])
            if ok[r]:
                g.append(int(bits[r]))
    assert res[0][1].tolist() == list(groups)
    for c, (kind, check) in enumerate([("min", False), ("max", False), ("min", True), ("max", True)], start=1):
        t, v, valid = res[c]
        assert t == T.DOUBLE and not res[c].approx
        for j, vals in enumerate(groups.values()):
            e = expect_grouped(vals, kind, check)
            assert bool(valid[j]) == (e is not None)
            if e is not None:
                assert int(v.view(np.uint64)[j]) == e, (kind, check, [hex(x) for x in vals])


@pytest.mark.parametrize("seed", range(6))
def test_global_float_min_max_bits(seed):
    rng = np.random.default_rng(100 + seed)
    batches, _ = _batches(rng, 4, 12, 1, null_frac=0.3)
    acts = [("action_min", [1]), ("action_max", [1]), ("action_min_true", [1]), ("action_max_true", [1])]
    res = oracle.aggregate(batches, [T.INT64, T.DOUBLE], [parse_action(n, i) for n, i in acts])
    per_batch = []
    for b in batches:
        _, bits, ok = _values(b)
        per_batch.append([int(bits[r]) for r in range(b.num_rows) if ok[r]])
    for c, (kind, check) in enumerate([("min", False), ("max", False), ("min", True), ("max", True)]):
        t, v, valid = res[c]
        e = expect_global(per_batch, kind, check)
        assert bool(valid[0]) == (e is not None)
        if e is not None:
            assert int(v.view(np.uint64)[0]) == e, (kind, check, [[hex(x) for x in bv] for bv in per_batch])


def test_global_differs_from_grouped_where_the_reference_does():
    """One batch [NaN_a, 1.0] then one [NaN_b, NaN_c, 2.0]: grouped plain min keeps
    NaN_a; the global min (LessNaNSafe always) is 1.0; global max with the NaN
    check is NaN_b (first NaN of the last NaN batch); grouped max is NaN_c."""
    a, bb, c = NANS[1], NANS[2], NANS[3]
    b1 = [a, 0x3FF0000000000000]
    b2 = [bb, c, 0x4000000000000000]

    def batch(vals, key):
        arr = np.array(vals, np.uint64).view(np.float64)
        return HostBatch(len(vals), [HostColumn(T.INT64, np.full(len(vals), key, np.int64)), HostColumn(T.DOUBLE, arr)])

    acts = [parse_action(n, i) for n, i in [("action_min", [1]), ("action_max_true", [1])]]
    g = oracle.aggregate([batch(b1, 0), batch(b2, 0)], [T.INT64, T.DOUBLE],
                         [parse_action("action_groupby", [0])] + acts)
    glob = oracle.aggregate([batch(b1, 0), batch(b2, 0)], [T.INT64, T.DOUBLE], acts)
    assert int(g[1][1].view(np.uint64)[0]) == a
    assert int(g[2][1].view(np.uint64)[0]) == c
    assert int(glob[0][1].view(np.uint64)[0]) == 0x3FF0000000000000
    assert int(glob[1][1].view(np.uint64)[0]) == bb


@pytest.mark.parametrize("kind", ["min", "max"])
def test_global_integer_all_null_batch_merges_zero(kind):
    """Batches [5, 7] then all-null: min -> 0 (the null scalar's value); max stays 7.
    An all-null batch first is ignored: [null, null] then [5, 7] -> 5 / 7.
    Negative values: [-5, -7] then all-null -> min -7, max 0."""
    def b(vals, nulls=False):
        n = len(vals)
        return HostBatch(n, [HostColumn(T.INT32, np.array(vals, np.int32),
                                        bitmap_from_bools(np.zeros(n, bool)) if nulls else None)])

    act = [parse_action(f"action_{kind}", [0])]
    got = lambda bs: oracle.aggregate(bs, [T.INT32], act)[0][1][0]  # noqa: E731
    assert got([b([5, 7]), b([1, 1], True)]) == (0 if kind == "min" else 7)
    assert got([b([1, 1], True), b([5, 7])]) == (5 if kind == "min" else 7)
    assert got([b([-5, -7]), b([1], True)]) == (-7 if kind == "min" else 0)
    assert got([b([5, 7]), b([3, 9])]) == (3 if kind == "min" else 9)


@pytest.mark.parametrize("ftype", [T.DOUBLE, T.FLOAT])
@pytest.mark.parametrize("seed", range(4))
def test_float_group_key_folds_zero_and_nan(ftype, seed):
    rng = np.random.default_rng(200 + seed)
    if ftype == T.DOUBLE:
        pool = np.array([POS0, NEG0, 0x3FF0000000000000] + NANS, np.uint64).view(np.float64)
        ubits = np.uint64
    else:
        pool = np.array([0, 0x80000000, 0x3F800000, 0x7FC00000, 0x7FC00ABC, 0xFFC00001, 0x7F800123],
                        np.uint32).view(np.float32)
        ubits = np.uint32
    batches = []
    for _ in range(3):
        n = 30
        k = pool[rng.integers(0, len(pool), n)]
        ok = rng.random(n) >= 0.15
        batches.append(HostBatch(n, [HostColumn(ftype, k, bitmap_from_bools(ok)),
                                     HostColumn(T.INT64, np.arange(n, dtype=np.int64))]))
    res = oracle.aggregate(batches, [ftype, T.INT64],
                           [parse_action(n, i) for n, i in [("action_groupby", [0]), ("action_count", [1])]])
    order, first_bits, counts = [], {}, {}
    from gazelle_plugin_amd.batch import bools_from_bitmap
    for b in batches:
        ok = bools_from_bitmap(b.columns[0].validity, b.num_rows)
        for r in range(b.num_rows):
            if not ok[r]:
                g = "null"
            else:
                x = float(b.columns[0].values[r])
                g = "nan" if math.isnan(x) else (0.0 if x == 0 else x)
            if g not in counts:
                order.append(g)
                first_bits[g] = None if g == "null" else int(b.columns[0].values[r:r + 1].view(ubits)[0])
                counts[g] = 0
            counts[g] += 1
    t, keys, valid = res[0]
    assert len(keys) == len(order)
    for j, g in enumerate(order):
        assert bool(valid[j]) == (g != "null")
        if g != "null":
            assert int(keys[j:j + 1].view(ubits)[0]) == first_bits[g]
        assert res[1][1][j] == counts[g]
