"""Helpers of the aggregate tests."""
import numpy as np

from gazelle_plugin_amd import types as T

CONFIG0_ACTIONS = [("action_groupby", [0]), ("action_sum", [1]), ("action_avg", [1]),
                   ("action_countLiteral_1", []), ("action_min", [2]), ("action_max", [2]),
                   ("action_sum", [2])]


def assert_results_equal(got, exp, rtol=1e-12):
    """got/exp: [(type, values, valid)] per result column, groups in the same order.

    Floating columns are compared by their bits (NaN payloads and the sign of a
    zero included), except the sums and averages that the expected column marks
    ``approx`` (the oracle's ``ResultColumn``): their summation order is free, so
    they agree within ``rtol``.  An expected column without that mark (a device
    result) is compared by bits where they agree and within ``rtol`` otherwise."""
    assert len(got) == len(exp), (len(got), len(exp))
    for c, (g, e) in enumerate(zip(got, exp)):
        tg, vg, okg = g
        te, ve, oke = e
        assert tg == te, f"type of column {c}: {tg} != {te}"
        np.testing.assert_array_equal(okg, oke, err_msg=f"validity of column {c}")
        vg, ve = np.asarray(vg)[oke], np.asarray(ve)[oke]
        if te in (T.DOUBLE, T.FLOAT):
            u = np.uint64 if te == T.DOUBLE else np.uint32
            same_bits = np.array_equal(vg.view(u), ve.view(u))
            approx = getattr(e, "approx", None)
            if approx is False or (approx is None and same_bits):
                bad = np.nonzero(vg.view(u) != ve.view(u))[0]
                assert same_bits, (f"column {c}: bits differ at {bad[:8].tolist()}: "
                                   f"{[hex(int(x)) for x in vg.view(u)[bad[:8]]]} != "
                                   f"{[hex(int(x)) for x in ve.view(u)[bad[:8]]]}")
                continue
            np.testing.assert_array_equal(np.isnan(vg), np.isnan(ve), err_msg=f"NaN of column {c}")
            m = ~np.isnan(ve)
            np.testing.assert_allclose(vg[m], ve[m], rtol=rtol, atol=0, err_msg=f"column {c}")
        else:
            np.testing.assert_array_equal(vg, ve, err_msg=f"column {c}")


def python_groupby(batches, key_col, cols):
    """Brute force: {key: [row values per col]} in first-seen order (None = null key)."""
    from gazelle_plugin_amd.batch import bools_from_bitmap

    groups = {}
    for b in batches:
        kv = bools_from_bitmap(b.columns[key_col].validity, b.num_rows)
        keys = b.columns[key_col].values
        vals = []
        for c in cols:
            col = b.columns[c]
            vals.append((col.values, bools_from_bitmap(col.validity, b.num_rows)))
        for r in range(b.num_rows):
            k = int(keys[r]) if kv[r] else None
            g = groups.setdefault(k, [[] for _ in cols])
            for j, (v, ok) in enumerate(vals):
                g[j].append(v[r].item() if ok[r] else None)
    return groups


def multikey_batches(sizes, seed=0, null_frac=0.15, domains=(5, 7)):
    """Batches of (int32 k0, int64 k1, double v, double k3, int16 v4): three key
    columns over small domains (k3 holds 0.0, -0.0 and two NaN payloads among
    its values) with nulls, so every null pattern and collision shows up."""
    from gazelle_plugin_amd.batch import HostBatch, HostColumn, bitmap_from_bools
    rng = np.random.default_rng(seed)
    k3_pool = np.array([0, 0x8000000000000000, 0x3FF8000000000000, 0x7FF8000000000000, 0x7FF8000000000001],
                       np.uint64).view(np.float64)  # 0.0, -0.0, 1.5, two NaN payloads
    out = []
    for n in sizes:
        def vb():
            return bitmap_from_bools(rng.random(n) >= null_frac) if n else None
        k0 = rng.integers(-2, domains[0] - 2, n).astype(np.int32)
        k1 = (rng.integers(0, domains[1], n) * 1000003).astype(np.int64)
        k3 = k3_pool[rng.integers(0, len(k3_pool), n)]
        v = rng.normal(size=n)
        v4 = rng.integers(-300, 300, n).astype(np.int16)
        out.append(HostBatch(n, [HostColumn(T.INT32, k0, vb()), HostColumn(T.INT64, k1, vb()),
                                 HostColumn(T.DOUBLE, v, vb()), HostColumn(T.DOUBLE, k3, vb()),
                                 HostColumn(T.INT16, v4, vb())]))
    return out


def python_groupby_keys(batches, key_cols, cols):
    """Brute force over several key columns: {key tuple: [row values per col]} in
    first-seen order, where a key tuple holds None for a null key and the raw
    bits of the value otherwise (so -0.0 and 0.0 differ, NaN payloads differ)."""
    from gazelle_plugin_amd.batch import bools_from_bitmap

    groups = {}
    for b in batches:
        keys = []
        for c in key_cols:
            col = b.columns[c]
            raw = col.values[: b.num_rows].view(np.dtype(f"u{col.values.dtype.itemsize}"))
            keys.append((raw, bools_from_bitmap(col.validity, b.num_rows)))
        vals = []
        for c in cols:
            col = b.columns[c]
            vals.append((col.values, bools_from_bitmap(col.validity, b.num_rows)))
        for r in range(b.num_rows):
            k = tuple(int(raw[r]) if ok[r] else None for raw, ok in keys)
            g = groups.setdefault(k, [[] for _ in cols])
            for j, (v, ok) in enumerate(vals):
                g[j].append(v[r].item() if ok[r] else None)
    return groups
