"""The partition writer's buffer policy on the GPU against the oracle's batch
by batch restatement of the reference (Splitter::DoSplit): record batch cuts
with buffer_size, prefer_spill and offheap_per_task (CalculateSplitBatchSize),
RawPartitionLengths with the in-memory validity rules, the LZ4_FRAME stream's
uncompressed record batches of at most batch_compress_threshold rows, and the
SplitResult metrics."""
import numpy as np
import pyarrow as pa
import pytest

import oracle
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.batch import DeviceBatch, bitmap_from_bools
from gazelle_plugin_amd.splitter import SplitOptions, Splitter
from splitcheck import random_batches

pytestmark = pytest.mark.gpu

ARROW = {T.INT64: pa.int64(), T.INT32: pa.int32(), T.DOUBLE: pa.float64(), T.STRING: pa.string(),
         T.BOOL: pa.bool_(), T.NA: pa.null(), T.DATE32: pa.date32()}


def _batches(types, sizes, seed, null_batches):
    """Random batches; only the batches listed in null_batches hold nulls (so
    the sticky and lazy validity rules matter)."""
    bs = random_batches(types, sizes, seed=seed, null_frac=0.2)
    for i, b in enumerate(bs):
        if i not in null_batches:
            for c in b.columns:
                if c.validity is not None:
                    c.validity = bitmap_from_bools(np.ones(b.num_rows, bool))
    return bs


def _run(ctx, batches, types, P, keys, opts, batched=True):
    sp = Splitter.Make(ctx, "hash", types, P, keys, opts)
    dev = [DeviceBatch.from_host(b) for b in batches]
    if batched:
        sp.SplitBatches(dev)
    else:
        for d in dev:
            sp.Split(d)
    sp.Stop()
    return sp


def _cuts(data, lengths):
    """Rows of every record batch of every partition in a data file."""
    out, pos = [], 0
    for n in lengths:
        n = int(n)
        rows = [] if n == 0 else [rb.num_rows for rb in pa.ipc.open_stream(pa.py_buffer(data[pos:pos + n].tobytes()))]
        out.append(rows)
        pos += n
    return out


CASES = [
    dict(buffer_size=4096, offheap_per_task=0, prefer_spill=True),
    dict(buffer_size=300, offheap_per_task=0, prefer_spill=True),
    dict(buffer_size=300, offheap_per_task=0, prefer_spill=False),
    dict(buffer_size=4096, offheap_per_task=1 << 20, prefer_spill=True),
    dict(buffer_size=4096, offheap_per_task=150_000, prefer_spill=False),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("batched", [True, False])
def test_cuts_and_raw_lengths_match_the_reference_policy(ctx, case, batched):
    kw = CASES[case]
    types = [T.INT64, T.STRING, T.DOUBLE, T.BOOL, T.INT32, T.NA, T.DATE32]
    batches = _batches(types, [3000, 1, 2500, 4096, 777, 2000, 3333], seed=31 + case, null_batches={2, 5})
    for b in batches:  # non-empty strings in every batch (the offheap row size reads the first)
        off = b.columns[1].offsets
        lens = np.maximum(np.diff(off), 1)
        b.columns[1].offsets = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        b.columns[1].values = np.resize(b.columns[1].values, int(b.columns[1].offsets[-1])).astype(np.uint8)
    P = 13
    pids = [oracle.split_pids([b], types, P, [0, 1]) for b in batches]
    exp = oracle.split_buffer_policy(batches, types, pids, P, **kw)
    sp = _run(ctx, batches, types, P, [0, 1], SplitOptions(**kw), batched)
    schema = pa.schema([(f"c{i}", ARROW[t]) for i, t in enumerate(types)])
    data, lengths = sp.ipc_stream(schema.serialize().to_pybytes())
    assert sp.BatchRows() == exp["batch_rows"]
    assert _cuts(data, lengths) == exp["cuts"]
    np.testing.assert_array_equal(sp.RawPartitionLengths(), exp["raw"])
    m = sp.Metrics()
    assert m["total_bytes_written"] == len(data) and m["total_spill_time"] == 0 and m["total_bytes_spilled"] == 0
    assert m["total_compute_pid_time"] > 0 and m["total_write_time"] > 0
    sp.close()


def _batch_compression(data, lengths, schema):
    """(rows, compressed?) of every record batch message of a data file: an LZ4
    body starts with its first buffer's int64 uncompressed length and the frame
    magic."""
    out, pos = [], 0
    for n in lengths:
        n = int(n)
        if n:
            for msg in pa.ipc.MessageReader.open_stream(pa.py_buffer(data[pos:pos + n].tobytes())):
                if msg.type != "record batch":
                    continue
                rb = pa.ipc.read_record_batch(msg, schema)
                out.append((rb.num_rows, msg.body.to_pybytes()[8:12] == b"\x04\x22\x4d\x18"))
        pos += n
    return out


@pytest.mark.parametrize("threshold", [256, 0, 10_000])
def test_lz4_batches_at_or_below_the_threshold_stay_uncompressed(ctx, threshold):
    types = [T.INT64, T.STRING, T.INT32]
    batches = random_batches(types, [4000, 300, 100, 2000, 50], seed=41, null_frac=0.1)
    P = 40  # about 100 rows per partition and input batch: cached batches of 300-600 rows, small last ones
    opts = SplitOptions(buffer_size=300, batch_compress_threshold=threshold)
    sp = Splitter.Make(ctx, "hash", types, P, [0], opts)
    sp.SetCompressType("lz4")
    sp.SplitBatches([DeviceBatch.from_host(b) for b in batches])
    sp.Stop()
    schema = pa.schema([(f"c{i}", ARROW[t]) for i, t in enumerate(types)])
    data, lengths = sp.ipc_stream(schema.serialize().to_pybytes())
    sp.close()
    seen = _batch_compression(data, lengths, schema)
    if threshold == 256:
        assert any(r <= threshold for r, _ in seen) and any(r > threshold for r, _ in seen)
    for rows, compressed in seen:
        assert compressed == (rows > threshold), (rows, threshold)
    # every partition still decodes to the oracle's rows
    exp = oracle.split(batches, types, P, [0])
    pos = 0
    for p, n in enumerate(lengths):
        n = int(n)
        got = sum(rb.num_rows for rb in pa.ipc.open_stream(pa.py_buffer(data[pos:pos + n].tobytes()))) if n else 0
        assert got == int(exp["row_start"][p + 1] - exp["row_start"][p])
        pos += n
