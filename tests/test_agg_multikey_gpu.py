"""Parity of the HIP aggregate grouped by several key columns (the reference's
UnsafeRow key, hash_aggregate_kernel.cc:970-1000) with the oracle through the
C ABI: groups in first-seen order, key columns bit-exact (raw bits, so -0.0,
0.0 and NaN payloads are told apart), integer results bit-exact, FP64 results
within 1e-12 relative."""
import ctypes

import numpy as np
import pytest

import oracle
from gazelle_plugin_amd import _lib as L
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.aggregate import AggOptions, HashAggregateKernel, parse_action
from gazelle_plugin_amd.batch import DeviceBatch, HostBatch, HostColumn, bitmap_from_bools
from aggcheck import assert_results_equal, multikey_batches

pytestmark = pytest.mark.gpu

RTOL = 1e-12
MK_TYPES = [T.INT32, T.INT64, T.DOUBLE, T.DOUBLE, T.INT16]
MK_ACTIONS = [("action_groupby", [0]), ("action_sum", [2]), ("action_groupby", [1]),
              ("action_countLiteral_1", []), ("action_groupby", [3]), ("action_min", [4]),
              ("action_max", [4]), ("action_sum_count", [2]), ("action_count", [4]),
              ("action_avg", [2]), ("action_max_true", [2])]


def run_gpu(ctx, batches, types, actions, mode="batched", **opts):
    k = HashAggregateKernel.Make(ctx, types, actions, AggOptions(**opts))
    it = k.MakeResultIterator()
    dev = [DeviceBatch.from_host(b) for b in batches]
    if mode == "batched":
        it.ProcessBatches(dev)
    elif mode == "one":
        for d in dev:
            it.ProcessAndCacheOne(d)
    else:  # a few batches per call
        for i in range(0, len(dev), 3):
            it.ProcessBatches(dev[i:i + 3])
    res = it.results()
    it.close()
    return res


def expected(batches, types, actions):
    return oracle.aggregate(batches, types, [parse_action(n, i) for n, i in actions])


def _assert_keys_bits(got, exp, cols):
    for c in cols:
        (tg, vg, okg), (te, ve, oke) = got[c], exp[c]
        np.testing.assert_array_equal(okg, oke, err_msg=f"validity of key column {c}")
        u = np.dtype(f"u{ve.dtype.itemsize}")
        np.testing.assert_array_equal(vg.view(u)[oke], ve.view(u)[oke], err_msg=f"bits of key column {c}")


@pytest.mark.parametrize("mode", ["batched", "one", "calls"])
def test_multi_key_small_domains(ctx, mode):
    batches = multikey_batches([3000, 0, 1, 2500, 4096, 17], seed=11)
    got = run_gpu(ctx, batches, MK_TYPES, MK_ACTIONS, mode)
    exp = expected(batches, MK_TYPES, MK_ACTIONS)
    _assert_keys_bits(got, exp, [0, 2, 4])
    assert_results_equal(got, exp, RTOL)


@pytest.mark.parametrize("placement", [1, 2, 3])
def test_multi_key_many_groups(ctx, placement):
    """~1e5 distinct key tuples over 300k rows in 4096-row batches."""
    batches = multikey_batches([4096] * 73 + [1000], seed=5, null_frac=0.02, domains=(2000, 60))
    got = run_gpu(ctx, batches, MK_TYPES, MK_ACTIONS, "batched", placement=placement)
    exp = expected(batches, MK_TYPES, MK_ACTIONS)
    assert len(exp[0][1]) > 50000
    _assert_keys_bits(got, exp, [0, 2, 4])
    assert_results_equal(got, exp, RTOL)


def test_multi_key_dictionary_grows_across_calls(ctx):
    """Calls that keep adding new key tuples: the code store and slot table
    grow and re-insert earlier codes between calls."""
    rng = np.random.default_rng(9)
    batches = []
    for i in range(12):
        n = 5000
        a = (rng.integers(0, 40000, n) + i * 40000).astype(np.int64)
        b = rng.integers(0, 3, n).astype(np.int8)
        v = rng.integers(-1000, 1000, n).astype(np.int64)
        batches.append(HostBatch(n, [HostColumn(T.INT64, a), HostColumn(T.INT8, b, bitmap_from_bools(rng.random(n) > 0.1)),
                                     HostColumn(T.INT64, v)]))
    types = [T.INT64, T.INT8, T.INT64]
    acts = [("action_groupby", [1]), ("action_groupby", [0]), ("action_sum", [2]), ("action_countLiteral_1", [])]
    got = run_gpu(ctx, batches, types, acts, "one")
    exp = expected(batches, types, acts)
    _assert_keys_bits(got, exp, [0, 1])
    assert_results_equal(got, exp, RTOL)


def test_multi_key_types_and_equality_cases(ctx):
    """Every fixed-width key width, and the UnsafeRow equality cases: the value
    under a null is ignored, (null, 5) differs from (5, null), -0.0 from 0.0,
    NaN payloads from each other."""
    k0 = np.array([5, 9, 5, 5, 7, 7, 7, 7, 7, 7], np.int64)
    k0_ok = bitmap_from_bools(np.array([0, 0, 1, 1, 1, 1, 1, 1, 1, 1], bool))
    k1 = np.array([5, 5, 3, 3, 0, 0, 0, 0, 0, 0], np.int64)
    k1_ok = bitmap_from_bools(np.array([1, 1, 0, 0, 1, 1, 1, 1, 1, 1], bool))
    d = np.array([0, 0, 0, 0, 0, 0x8000000000000000, 0, 0x7FF8000000000000, 0x7FF8000000000001, 0],
                 np.uint64).view(np.float64)
    f = np.array([1, 1, 1, 1, 1, 1, 1, 1, 1, -0.0], np.float32)
    u8 = np.full(10, 200, np.uint8)
    i16 = np.full(10, -3, np.int16)
    d32 = np.full(10, 19000, np.int32)
    ts = np.full(10, 1 << 40, np.int64)
    b = HostBatch(10, [HostColumn(T.INT64, k0, k0_ok), HostColumn(T.INT64, k1, k1_ok), HostColumn(T.DOUBLE, d),
                       HostColumn(T.FLOAT, f), HostColumn(T.UINT8, u8), HostColumn(T.INT16, i16),
                       HostColumn(T.DATE32, d32), HostColumn(T.TIMESTAMP, ts)])
    types = [T.INT64, T.INT64, T.DOUBLE, T.FLOAT, T.UINT8, T.INT16, T.DATE32, T.TIMESTAMP]
    acts = [("action_groupby", [c]) for c in range(8)] + [("action_countLiteral_1", [])]
    got = run_gpu(ctx, [b], types, acts)
    exp = expected([b], types, acts)
    assert exp[8][1].tolist() == [2, 2, 2, 1, 1, 1, 1]
    _assert_keys_bits(got, exp, list(range(8)))
    assert_results_equal(got, exp, RTOL)


def test_multi_key_rejects_what_it_does_not_group(ctx):
    """Two utf8 keys (their bytes meet in the UnsafeRow with no length, so
    ("ab", "c") and ("a", "bc") are one group there) and the set/planned forms
    of a multi-key plan are NotImplemented."""
    h = ctypes.c_void_p()
    acts = (L.GzAggAction * 2)()
    acts[0].kind, acts[0].input0, acts[0].input1 = 0, 0, -1
    acts[1].kind, acts[1].input0, acts[1].input1 = 0, 1, -1
    types = L.i32_array([T.STRING, T.STRING])
    assert L.lib().gz_agg_make(ctx.handle, 2, types, 2, acts, None, ctypes.byref(h)) == 10
    types = L.i32_array([T.INT64, T.INT32])
    assert L.lib().gz_agg_make(ctx.handle, 2, types, 2, acts, None, ctypes.byref(h)) == 0
    try:
        out = ctypes.c_int32()
        assert L.lib().gz_agg_set_placement(h, 0, ctypes.byref(out)) == 10
    finally:
        L.lib().gz_agg_release(h)


def _utf8_batches(sizes, seed, pool_size=300, null_frac=0.1):
    """(int32 k, utf8 s, double v) batches; s from a pool of random byte strings
    of length 0-40 plus a few of 100-300 bytes (the empty string included)."""
    rng = np.random.default_rng(seed)
    pool = [b""] + [rng.bytes(int(rng.integers(1, 41))) for _ in range(pool_size - 6)]
    pool += [rng.bytes(int(rng.integers(100, 301))) for _ in range(5)]
    out = []
    for n in sizes:
        pick = rng.integers(0, len(pool), n)
        enc = [pool[i] for i in pick]
        off = np.zeros(n + 1, np.int32)
        off[1:] = np.cumsum([len(e) for e in enc])
        data = np.frombuffer(b"".join(enc), np.uint8).copy() if off[-1] else np.zeros(0, np.uint8)
        sv = bitmap_from_bools(rng.random(n) >= null_frac)
        k = rng.integers(0, 5, n).astype(np.int32)
        v = rng.normal(size=n)
        out.append(HostBatch(n, [HostColumn(T.INT32, k, bitmap_from_bools(rng.random(n) >= null_frac)),
                                 HostColumn(T.STRING, data, sv, off), HostColumn(T.DOUBLE, v)]))
    return out


UTF8_TYPES = [T.INT32, T.STRING, T.DOUBLE]


@pytest.mark.parametrize("mode", ["batched", "one", "calls"])
def test_utf8_key_alone(ctx, mode):
    """One utf8 key (the reference's LargeStringHashMap route): groups by the
    bytes, the empty string and the null key are groups of their own."""
    batches = _utf8_batches([3000, 0, 1, 2500, 4096], seed=21)
    acts = [("action_sum", [2]), ("action_groupby", [1]), ("action_countLiteral_1", []), ("action_min", [0]),
            ("action_count", [2])]
    got = run_gpu(ctx, batches, UTF8_TYPES, acts, mode)
    exp = expected(batches, UTF8_TYPES, acts)
    assert len(exp[1][1]) > 250
    assert_results_equal(got, exp, RTOL)


@pytest.mark.parametrize("mode", ["batched", "one", "calls"])
def test_utf8_then_int_keys(ctx, mode):
    """(utf8, int32) keys: the utf8 key first in the UnsafeRow.  With equal
    null and empty masks, equal row bytes need equal utf8 lengths, so groups
    are field by field; the key columns come back in the plan's order."""
    batches = _utf8_batches([4096] * 4 + [77, 0, 1], seed=23)
    acts = [("action_groupby", [1]), ("action_sum", [2]), ("action_groupby", [0]), ("action_countLiteral_1", []),
            ("action_min", [2])]
    got = run_gpu(ctx, batches, UTF8_TYPES, acts, mode)
    exp = expected(batches, UTF8_TYPES, acts)
    assert len(exp[0][1]) > 1000
    _assert_keys_bits(got, exp, [2])
    assert_results_equal(got, exp, RTOL)


@pytest.mark.parametrize("mode", ["batched", "one"])
def test_int_then_utf8_keys(ctx, mode):
    """(int32, utf8) keys: the utf8 key last in the UnsafeRow, ~1500 groups,
    strings up to 300 bytes, the byte heap growing across calls."""
    batches = _utf8_batches([4096] * 9 + [77], seed=22)
    acts = [("action_groupby", [0]), ("action_groupby", [1]), ("action_sum", [2]), ("action_countLiteral_1", []),
            ("action_max", [2])]
    got = run_gpu(ctx, batches, UTF8_TYPES, acts, mode)
    exp = expected(batches, UTF8_TYPES, acts)
    assert len(exp[0][1]) > 1000
    _assert_keys_bits(got, exp, [0])
    assert_results_equal(got, exp, RTOL)
