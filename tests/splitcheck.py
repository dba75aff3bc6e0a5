"""Helpers shared by the split tests: synthetic schemas and layout comparison."""
from __future__ import annotations

import numpy as np

from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.batch import HostBatch, HostColumn, bitmap_from_bools


def random_column(rng, t: int, n: int, null_frac: float = 0.0) -> HostColumn:
    valid = rng.random(n) >= null_frac if null_frac > 0 else np.ones(n, bool)
    vb = bitmap_from_bools(valid) if null_frac > 0 else None
    if t == T.NA:
        return HostColumn(t, None, None, None)
    if t == T.BOOL:
        return HostColumn(t, bitmap_from_bools(rng.random(n) < 0.5), vb)
    if T.is_binary(t):
        lens = rng.integers(0, 40, size=n).astype(np.int32)
        off = np.zeros(n + 1, np.int32)
        off[1:] = np.cumsum(lens)
        data = rng.integers(0, 256, size=int(off[-1]), dtype=np.uint8)
        return HostColumn(t, data, vb, off)
    w = T.width(t)
    raw = rng.integers(0, 256, size=n * w, dtype=np.uint8)
    return HostColumn(t, raw.view(T.numpy_dtype(t)) if t != T.DECIMAL128 else raw, vb)


def ordinary_float_keys(batches, cols):
    """Replace -0.0 and NaN bit patterns in float / double key columns (whose
    Spark hash the split refuses) by other values, in place."""
    for b in batches:
        for c in cols:
            col = b.columns[c]
            if col.type_id == T.DOUBLE:
                u, sign, exp = col.values.view(np.uint64), np.uint64(1 << 63), np.uint64(0x7FF0000000000000)
            elif col.type_id == T.FLOAT:
                u, sign, exp = col.values.view(np.uint32), np.uint32(1 << 31), np.uint32(0x7F800000)
            else:
                continue
            u[(u & exp) == exp] ^= exp & (exp >> u.dtype.type(1))  # drop the exponent's top bit: finite
            u[u == sign] = 0
    return batches


def random_batches(types, sizes, seed=0, null_frac=0.1):
    rng = np.random.default_rng(seed)
    return [HostBatch(n, [random_column(rng, t, n, null_frac if t != T.NA else 0) for t in types])
            for n in sizes]


def _u8(a):
    return np.ascontiguousarray(a).view(np.uint8).reshape(-1)


def assert_layout_equal(got, exp, types):
    """Bit-exact comparison of two concatenated partition layouts."""
    np.testing.assert_array_equal(got["row_start"], exp["row_start"], err_msg="row_start")
    N = int(exp["row_start"][-1])
    nbm = (N + 63) // 64 * 8
    for c, t in enumerate(types):
        g, e = got["columns"][c], exp["columns"][c]
        ev, gv = e.get("validity"), g.get("validity")
        if ev is not None or gv is not None:
            ones = np.full(nbm, 0xFF, np.uint8)
            ev = ones if ev is None else _u8(ev)[:nbm]
            gv = ones if gv is None else _u8(gv)[:nbm]
            np.testing.assert_array_equal(gv, ev, err_msg=f"validity of column {c}")
        if t == T.NA:
            continue
        if T.is_binary(t):
            np.testing.assert_array_equal(g["byte_start"], e["byte_start"], err_msg=f"byte_start {c}")
            np.testing.assert_array_equal(_u8(g["offsets"]), _u8(e["offsets"]), err_msg=f"offsets {c}")
            np.testing.assert_array_equal(_u8(g["data"]), _u8(e["data"]), err_msg=f"data {c}")
        elif t == T.BOOL:
            np.testing.assert_array_equal(_u8(g["values"])[:nbm], _u8(e["values"])[:nbm],
                                          err_msg=f"bool values {c}")
        else:
            w = T.width(t)
            np.testing.assert_array_equal(_u8(g["values"])[:N * w], _u8(e["values"])[:N * w],
                                          err_msg=f"values of column {c}")
