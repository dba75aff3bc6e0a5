"""Parity of the HIP hash join on several key columns (the reference's
UnsafeRow-keyed HashRelation and probe) with the oracle through the C ABI:
output rows bit-exact, in probe order with matches in build insertion order."""
import ctypes

import numpy as np
import pytest

import oracle
from gazelle_plugin_amd import _lib as L
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.batch import HostBatch, HostColumn
from test_join_gpu import check_output, host_values, run_case
from test_oracle_join import JT, multikey_join_batches

pytestmark = pytest.mark.gpu

BTYPES = [T.INT32, T.INT64, T.DOUBLE]
PTYPES = [T.INT64, T.INT32, T.INT32]
RESULT = [(1, 1), (1, 2), (0, 2), (0, 0)]


def expected(build, probe, bkeys, pkeys, join_type, result, builder_type=1):
    bflat = [(bi, r) for bi, b in enumerate(build) for r in range(b.num_rows)]
    outs = oracle.join_multi(build, bkeys, probe, pkeys, JT[join_type], builder_type)
    res = []
    for pb, out in zip(probe, outs):
        prow, brow = out if join_type == "inner" else (out, None)
        cols = []
        for s, i in result:
            if s == 1:
                cols.append(host_values(pb, i, list(prow)))
            elif brow is None:
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


@pytest.mark.parametrize("join_type", ["inner", "semi", "anti"])
@pytest.mark.parametrize("batched", [True, False])
def test_two_key_join(ctx, join_type, batched):
    build, probe = multikey_join_batches(7)
    result = RESULT if join_type == "inner" else [(1, 1), (1, 2), (1, 0)]
    got = run_case(ctx, build, probe, BTYPES, PTYPES, [0, 1], [2, 0], join_type, False, result, batched)
    exp = expected(build, probe, [0, 1], [2, 0], join_type, result)
    for g, e in zip(got, exp):
        check_output(g, e)


def test_two_key_semi_skip_duplicates(ctx):
    build, probe = multikey_join_batches(8)
    result = [(1, 1), (1, 0)]
    got = run_case(ctx, build, probe, BTYPES, PTYPES, [0, 1], [2, 0], "semi", False, result, True, builder_type=3)
    exp = expected(build, probe, [0, 1], [2, 0], "semi", result, builder_type=3)
    for g, e in zip(got, exp):
        check_output(g, e)


def test_int_and_utf8_keys_join(ctx):
    """(int32, utf8) keys, the utf8 key last; strings with a shared prefix,
    the empty string and nulls."""
    rng = np.random.default_rng(3)
    pool = [b"", b"a", b"ab", b"abc", b"b" * 70, b"\x00"]

    def batch(n, extra):
        from gazelle_plugin_amd.batch import bitmap_from_bools
        enc = [pool[i] for i in rng.integers(0, len(pool), n)]
        off = np.zeros(n + 1, np.int32)
        off[1:] = np.cumsum([len(e) for e in enc])
        data = np.frombuffer(b"".join(enc), np.uint8).copy() if off[-1] else np.zeros(0, np.uint8)
        return HostBatch(n, [HostColumn(T.INT32, rng.integers(0, 3, n).astype(np.int32),
                                        bitmap_from_bools(rng.random(n) > 0.1)),
                             HostColumn(T.STRING, data, bitmap_from_bools(rng.random(n) > 0.1), off),
                             HostColumn(T.INT64, rng.integers(0, 1000, n).astype(np.int64) + extra)])

    build = [batch(200, 0), batch(55, 10000)]
    probe = [batch(300, 20000), batch(1, 30000)]
    types = [T.INT32, T.STRING, T.INT64]
    result = [(1, 2), (0, 2), (0, 1)]
    got = run_case(ctx, build, probe, types, types, [0, 1], [0, 1], "inner", False, result, True)
    exp = expected(build, probe, [0, 1], [0, 1], "inner", result)
    for g, e in zip(got, exp):
        check_output(g, e)


def test_multi_key_join_rejects(ctx):
    """Key types that differ between the sides, and a null-aware anti join
    with several keys."""
    from gazelle_plugin_amd.join import HashRelationKernel
    from gazelle_plugin_amd.batch import DeviceBatch
    build, _ = multikey_join_batches(9)
    rel = HashRelationKernel.Make(ctx, BTYPES, [0, 1])
    rel.EvaluateBatches([DeviceBatch.from_host(b) for b in build])
    rel.Finish()
    h = ctypes.c_void_p()
    sides, idx = L.i32_array([1]), L.i32_array([0])
    rc = L.lib().gz_probe_make_keys(rel.handle, 3, L.i32_array([T.INT64, T.INT64, T.INT64]), 2,
                                    L.i32_array([2, 0]), 0, 0, 1, sides, idx, ctypes.byref(h))
    assert rc == 3, rc  # GZ_TYPE_ERROR
    rc = L.lib().gz_probe_make_keys(rel.handle, 3, L.i32_array(PTYPES), 2, L.i32_array([2, 0]), 2, 1, 1, sides,
                                    idx, ctypes.byref(h))
    assert rc == 10, rc  # GZ_NOT_IMPLEMENTED
    rel.close()
