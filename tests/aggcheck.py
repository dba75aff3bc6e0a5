"""Helpers of the aggregate tests."""
import numpy as np

from gazelle_plugin_amd import types as T

CONFIG0_ACTIONS = [("action_groupby", [0]), ("action_sum", [1]), ("action_avg", [1]),
                   ("action_countLiteral_1", []), ("action_min", [2]), ("action_max", [2]),
                   ("action_sum", [2])]


def assert_results_equal(got, exp, rtol=1e-12):
    """got/exp: [(type, values, valid)] per result column, groups in the same order."""
    assert len(got) == len(exp), (len(got), len(exp))
    for c, ((tg, vg, okg), (te, ve, oke)) in enumerate(zip(got, exp)):
        assert tg == te, f"type of column {c}: {tg} != {te}"
        np.testing.assert_array_equal(okg, oke, err_msg=f"validity of column {c}")
        vg, ve = vg[oke], ve[oke]
        if te in (T.DOUBLE, T.FLOAT):
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
