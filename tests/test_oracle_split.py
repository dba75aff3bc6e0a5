"""Pins of the oracle's split against the reference's own Splitter tests
(shuffle_split_test.cc) and against brute force on small random inputs."""
import numpy as np
import pytest

import oracle
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.batch import HostBatch, HostColumn, bools_from_bitmap
from gazelle_plugin_amd.partition_writer import reference_batch_rows
from golden import fixtures as F
from splitcheck import random_batches


def decode(layout, c, p):
    """Python values of output column c, partition p (None = null)."""
    rs = layout["row_start"]
    a, b = int(rs[p]), int(rs[p + 1])
    col = layout["columns"][c]
    t = col["type"]
    n = b - a
    valid = np.ones(n, bool) if col.get("validity") is None else \
        bools_from_bitmap(col["validity"], int(rs[-1]))[a:b]
    if t == T.NA:
        return [None] * n
    if t == T.BOOL:
        vals = list(bools_from_bitmap(col["values"], int(rs[-1]))[a:b])
    elif T.is_binary(t):
        off = col["offsets"][a + p: b + p + 1]
        base = int(col["byte_start"][p])
        vals = [bytes(col["data"][base + off[i]: base + off[i + 1]]) for i in range(n)]
    else:
        w = T.width(t)
        raw = np.asarray(col["values"]).view(np.uint8)[a * w: b * w]
        vals = [bytes(raw[i * w:(i + 1) * w]) for i in range(n)]
    return [v if ok else None for v, ok in zip(vals, valid)]


def decode_rows(batch: HostBatch, c, rows):
    col = batch.columns[c]
    t = col.type_id
    valid = bools_from_bitmap(col.validity, batch.num_rows)
    out = []
    for r in rows:
        if not valid[r] or t == T.NA:
            out.append(None)
        elif t == T.BOOL:
            out.append(bool(bools_from_bitmap(col.values, batch.num_rows)[r]))
        elif T.is_binary(t):
            out.append(bytes(col.values[col.offsets[r]:col.offsets[r + 1]]))
        else:
            w = T.width(t)
            raw = np.asarray(col.values).view(np.uint8)
            out.append(bytes(raw[r * w:(r + 1) * w]))
    return out


@pytest.mark.parametrize("case", [0, 1, 2])
def test_reference_splitter_fixtures(case):
    fx = F.load("split_fixtures.json")
    test = fx["tests"][case]
    schema = fx["schema"]
    named = {k: F.batch(schema, v) for k, v in fx["batches"].items()}
    inputs = [named[k] for k in test["split"]]
    types = [F.type_id(t) for _, t in schema]
    mode = test["short_name"]
    if mode == "range":
        inputs = [HostBatch(b.num_rows, [HostColumn(T.INT32, np.array(test["pids"][k], np.int32))] + b.columns)
                  for b, k in zip(inputs, test["split"])]
        types = [T.INT32] + types
    P = test["num_partitions"]
    lay = oracle.split(inputs, types, P, [], mode=mode)
    out_types = types[1:] if mode == "range" else types
    for p, parts in enumerate(test["expected"]):
        for c in range(len(out_types)):
            exp = []
            for name, rows in parts:
                exp += decode_rows(named[name], c, rows)
            assert decode(lay, c, p) == exp, (test["name"], p, c)
    # output record batches per partition, as the reference's buffer policy cuts them
    counts = np.stack([np.bincount(oracle.split_pids([b], types, P, [], mode=mode), minlength=P)
                       for b in inputs]) if mode != "range" else \
        np.stack([np.bincount(np.array(test["pids"][k]), minlength=P) for k in test["split"]])
    for p, parts in enumerate(test["expected"]):
        assert reference_batch_rows(counts[:, p], test["buffer_size"]) == [len(r) for _, r in parts]


def test_random_split_is_a_stable_partition():
    types = [T.INT64, T.INT32, T.DOUBLE, T.STRING, T.BOOL, T.INT8, T.DECIMAL128]
    batches = random_batches(types, [100, 0, 257, 1, 64], seed=3, null_frac=0.2)
    P = 7
    lay = oracle.split(batches, types, P, [0, 3])
    pids = lay["pids"]
    # brute force: rows of partition p are the input rows with pid p, in input order
    flat = [(bi, r) for bi, b in enumerate(batches) for r in range(b.num_rows)]
    for c in range(len(types)):
        allvals = [decode_rows(batches[bi], c, [r])[0] for bi, r in flat]
        for p in range(P):
            exp = [v for v, q in zip(allvals, pids) if q == p]
            assert decode(lay, c, p) == exp, (c, p)
    assert int(lay["row_start"][-1]) == sum(b.num_rows for b in batches)


def test_range_pid_out_of_range_is_invalid():
    b = HostBatch(3, [HostColumn(T.INT32, np.array([0, 2, 1], np.int32)), HostColumn(T.INT64, np.zeros(3, np.int64))])
    with pytest.raises(ValueError):
        oracle.split([b], [T.INT32, T.INT64], 2, [], mode="range")
