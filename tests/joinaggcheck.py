"""Helpers of the fused join -> aggregate tests: the expected result is the
oracle's aggregate over the oracle's inner join output, materialized in the
reference's order (probe rows in order, matches in build insertion order)."""
import numpy as np

import oracle
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.aggregate import parse_action
from gazelle_plugin_amd.batch import HostBatch, HostColumn, bitmap_from_bools, bools_from_bitmap


def key_of(batch, c):
    col = batch.columns[c]
    return (col.values, col.validity, batch.num_rows)


def joined_rows(build, probe, bkey, pkey, key_type, builder_type=1):
    """Per probe batch: (probe rows, global build rows) of the inner join."""
    return oracle.join([key_of(b, bkey) for b in build], [key_of(b, pkey) for b in probe], key_type, 0,
                       False, builder_type)


def _gather(batches, c, rows_per_batch, global_build):
    """Values and validity of column c at the given rows."""
    vals, valid = [], []
    if global_build is not None:
        starts = np.cumsum([0] + [b.num_rows for b in batches])
        rows = global_build
        bi = np.searchsorted(starts, rows, side="right") - 1
        for b in range(len(batches)):
            sel = rows[bi == b] - starts[b]
            col = batches[b].columns[c]
            # keep the order of `rows`: fill by position
            vals.append((np.nonzero(bi == b)[0], np.asarray(col.values)[sel],
                         bools_from_bitmap(col.validity, batches[b].num_rows)[sel]))
        n = len(rows)
        t = batches[0].columns[c].type_id
        out_v = np.zeros(n, T.numpy_dtype(t))
        out_ok = np.ones(n, bool)
        for pos, v, ok in vals:
            out_v[pos] = v
            out_ok[pos] = ok
        return out_v, out_ok
    outs_v, outs_ok = [], []
    for b, rows in zip(batches, rows_per_batch):
        col = b.columns[c]
        outs_v.append(np.asarray(col.values)[rows])
        outs_ok.append(bools_from_bitmap(col.validity, b.num_rows)[rows])
    return np.concatenate(outs_v), np.concatenate(outs_ok)


def materialize(build, probe, bkey, pkey, key_type, inputs, builder_type=1) -> HostBatch:
    """The join output restricted to the aggregate's input columns [(side, col)]."""
    outs = joined_rows(build, probe, bkey, pkey, key_type, builder_type)
    prow = [np.asarray(p, np.int64) for p, _ in outs]
    brow = np.concatenate([np.asarray(b, np.int64) for _, b in outs]) if outs else np.zeros(0, np.int64)
    cols = []
    for side, c in inputs:
        if side == 0:
            v, ok = _gather(build, c, None, brow)
            t = build[0].columns[c].type_id
        else:
            v, ok = _gather(probe, c, prow, None)
            t = probe[0].columns[c].type_id
        cols.append(HostColumn(t, v, None if ok.all() else bitmap_from_bools(ok)))
    return HostBatch(len(brow), cols)


def expected(build, probe, bkey, pkey, key_type, inputs, actions, builder_type=1):
    hb = materialize(build, probe, bkey, pkey, key_type, inputs, builder_type)
    types = [c.type_id for c in hb.columns]
    acts = [parse_action(n, i) for n, i in actions]
    return oracle.aggregate([hb] if hb.num_rows else [], types, acts)
