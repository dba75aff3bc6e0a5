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


# ---- the reference's partition buffers: record batch cuts and RawPartitionLengths

def _b(cols):
    """cols: [(type, values, nulls mask or None)]; utf8 values are str / None."""
    from gazelle_plugin_amd.batch import HostBatch, HostColumn, bitmap_from_bools
    n = len(cols[0][1])
    out = []
    for t, vals, nulls in cols:
        vb = None if nulls is None else bitmap_from_bools(~np.asarray(nulls, bool))
        if t == T.STRING:
            enc = [(v or "").encode() for v in vals]
            off = np.zeros(n + 1, np.int32)
            off[1:] = np.cumsum([len(e) for e in enc])
            out.append(HostColumn(t, np.frombuffer(b"".join(enc), np.uint8).copy(), vb, off))
        else:
            out.append(HostColumn(t, np.asarray(vals, np.int64 if t == T.INT64 else np.int32), vb))
    return HostBatch(n, out)


def _policy(batches, types, **kw):
    pids = [np.zeros(b.num_rows, np.int32) for b in batches]
    return oracle.split_buffer_policy(batches, types, pids, 1, **kw)


def test_buffer_reuse_keeps_validity_and_reset_drops_it():
    """int64 a (never null), utf8 s; buffer_size 2, one partition.  Batch 0 has a
    null in s: s's buffer gets a validity bitmap lazily.  Batch 1 has none.
    Reused buffer (prefer_spill): both cached batches carry s's bitmap:
      cut 0: a 2*8 + s (1 + 3*4 + 1 byte) = 30; cut 1: a 16 + s (1 + 12 + 3) = 32.
    Reset buffer (no prefer_spill): cut 1 has no bitmap: 16 + 12 + 3 = 31."""
    types = [T.INT64, T.STRING]
    bs = [_b([(T.INT64, [1, 2], None), (T.STRING, ["x", None], [False, True])]),
          _b([(T.INT64, [3, 4], None), (T.STRING, ["yy", "z"], None)])]
    r = _policy(bs, types, buffer_size=2, prefer_spill=True)
    assert r["cuts"] == [[2, 2]] and r["raw"] == [30 + 32]
    r = _policy(bs, types, buffer_size=2, prefer_spill=False)
    assert r["cuts"] == [[2, 2]] and r["raw"] == [30 + 31]


def test_sticky_null_flag_compares_the_schema_index():
    """utf8 s, int64 a: the fixed-width slot of a is schema column 1, not below
    the fixed-width count (1), so a's null in batch 0 never sets its sticky flag;
    the reset buffer of batch 1 has no bitmap for a (splitter.cc:878-886):
      cut 0: s (12 + 2) + a (1 + 16) = 31; cut 1: s 14 + a 16 = 30."""
    types = [T.STRING, T.INT64]
    bs = [_b([(T.STRING, ["x", "y"], None), (T.INT64, [0, 1], [True, False])]),
          _b([(T.STRING, ["p", "q"], None), (T.INT64, [2, 3], None)])]
    r = _policy(bs, types, buffer_size=2, prefer_spill=False)
    assert r["raw"] == [31 + 30]
    # the other column order sets it: the reset buffer allocates a's bitmap (cut 1: 14 + 17)
    bs2 = [_b([(T.INT64, [0, 1], [True, False]), (T.STRING, ["x", "y"], None)]),
           _b([(T.INT64, [2, 3], None), (T.STRING, ["p", "q"], None)])]
    r = _policy(bs2, [T.INT64, T.STRING], buffer_size=2, prefer_spill=False)
    assert r["raw"] == [31 + 31]


def test_offheap_batch_size_from_the_first_batch():
    """CalculateSplitBatchSize: int64 + utf8 of 5 bytes per row in the first
    batch: 13 bytes per row; 65536 / 13 / 4 partitions >> 2 = 315 rows.  With
    utf8 first the int64's width is not counted (the schema index again): 5 bytes
    per row, 65536 / 5 / 4 >> 2 = 819."""
    n = 4096
    a = np.arange(n)
    s = ["abcde"] * n
    P = 4
    pids = [np.arange(n, dtype=np.int32) % P]
    r = oracle.split_buffer_policy([_b([(T.INT64, a, None), (T.STRING, s, None)])], [T.INT64, T.STRING], pids, P,
                                   buffer_size=4096, offheap_per_task=65536)
    assert r["batch_rows"] == 315
    r = oracle.split_buffer_policy([_b([(T.STRING, s, None), (T.INT64, a, None)])], [T.STRING, T.INT64], pids, P,
                                   buffer_size=4096, offheap_per_task=65536)
    assert r["batch_rows"] == 819
    r = oracle.split_buffer_policy([_b([(T.INT64, a, None), (T.STRING, s, None)])], [T.INT64, T.STRING], pids, P,
                                   buffer_size=200, offheap_per_task=65536)
    assert r["batch_rows"] == 200  # never above buffer_size


def test_cuts_follow_reference_batch_rows_without_nulls():
    """Null-free inputs: the cuts equal the buffer policy restated per partition
    (partition_writer.reference_batch_rows), and every cached batch's bytes are
    its rows' widths plus (n + 1) offsets and its string bytes."""
    from gazelle_plugin_amd.partition_writer import reference_batch_rows
    rng = np.random.default_rng(3)
    types = [T.INT64, T.STRING]
    bs = []
    for n in (700, 5, 1200, 64, 900):
        s = ["x" * int(k) for k in rng.integers(0, 9, n)]
        bs.append(_b([(T.INT64, rng.integers(0, 100, n), None), (T.STRING, s, None)]))
    P = 5
    pids = [rng.integers(0, P, b.num_rows).astype(np.int32) for b in bs]
    for prefer in (True, False):
        r = oracle.split_buffer_policy(bs, types, pids, P, buffer_size=150, prefer_spill=prefer)
        for p in range(P):
            counts = [int((pp == p).sum()) for pp in pids]
            assert r["cuts"][p] == reference_batch_rows(counts, 150, prefer)
            rows = sum(counts)
            nbytes = sum(int(np.diff(b.columns[1].offsets)[pp == p].sum()) for b, pp in zip(bs, pids))
            assert r["raw"][p] == rows * 8 + (rows + len(r["cuts"][p])) * 4 + nbytes
