"""The partition writer (Splitter::Stop's data file) on the GPU: every
partition's bytes are an Arrow IPC stream that pyarrow reads back, its record
batches cut as the reference's buffer policy cuts them, and its rows equal the
oracle's split, value for value."""
import numpy as np
import pyarrow as pa
import pytest

import oracle
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.batch import DeviceBatch
from gazelle_plugin_amd.partition_writer import reference_batch_rows
from gazelle_plugin_amd.splitter import SplitOptions, Splitter
from splitcheck import random_batches

pytestmark = pytest.mark.gpu

ARROW = {T.INT64: pa.int64(), T.INT32: pa.int32(), T.DOUBLE: pa.float64(), T.STRING: pa.string(),
         T.BOOL: pa.bool_(), T.INT8: pa.int8(), T.UINT16: pa.uint16(), T.NA: pa.null(), T.BINARY: pa.binary(),
         T.DATE32: pa.date32()}


def _bits(bm, start, n):
    return np.unpackbits(np.frombuffer(bm.tobytes(), np.uint8), bitorder="little")[start:start + n].astype(bool)


def _check_partition(arrays, exp, types, p):
    r0, r1 = int(exp["row_start"][p]), int(exp["row_start"][p + 1])
    n = r1 - r0
    for c, t in enumerate(types):
        a = arrays[c]
        assert len(a) == n
        col = exp["columns"][c]
        if t == T.NA:
            assert a.null_count == n
            continue
        valid = np.ones(n, bool) if col.get("validity") is None else _bits(col["validity"], r0, n)
        np.testing.assert_array_equal(np.asarray(a.is_valid()), valid, err_msg=f"validity p={p} c={c}")
        if T.is_binary(t):
            off = np.asarray(col["offsets"], np.int64)[r0 + p: r1 + p + 1]
            bs = int(col["byte_start"][p])
            data = np.asarray(col["data"], np.uint8)
            got = (a.cast(pa.binary()) if t == T.STRING else a).to_pylist()  # bytes, no utf8 decoding
            for i in range(n):
                if valid[i]:
                    assert got[i] == data[bs + off[i]: bs + off[i + 1]].tobytes(), f"bytes p={p} c={c} row {i}"
        elif t == T.BOOL:
            vals = np.asarray(a.fill_null(False), bool)
            np.testing.assert_array_equal(vals[valid], _bits(col["values"], r0, n)[valid])
        else:
            w = T.width(t)
            got = np.frombuffer(a.buffers()[1], np.uint8)[a.offset * w:(a.offset + n) * w]
            exp_v = np.asarray(col["values"]).view(np.uint8).reshape(-1)[r0 * w:r1 * w]
            np.testing.assert_array_equal(got, exp_v, err_msg=f"values p={p} c={c}")


# P = 2500: the splitter's sorted path (more than 2048 partitions), whose
# per-batch counts and buffer statistics the writer's cuts read
@pytest.mark.parametrize("batched,buffer_size,P", [(True, 4096, 11), (True, 300, 11), (False, 4096, 11),
                                                   (False, 300, 11), (True, 3, 2500)])
def test_data_file_decodes_to_the_split(ctx, batched, buffer_size, P):
    types = [T.INT64, T.INT32, T.DOUBLE, T.STRING, T.BOOL, T.INT8, T.UINT16, T.NA, T.BINARY, T.DATE32]
    batches = random_batches(types, [5000, 0, 1, 4096, 777, 3000], seed=21, null_frac=0.15)
    exp = oracle.split(batches, types, P, [0, 3])
    pids = oracle.split_pids(batches, types, P, [0, 3])
    sp = Splitter.Make(ctx, "hash", types, P, [0, 3], SplitOptions(buffer_size=buffer_size))
    dev = [DeviceBatch.from_host(b) for b in batches]
    if batched:
        sp.SplitBatches(dev)
    else:
        for d in dev:
            sp.Split(d)
    sp.Stop()
    schema = pa.schema([(f"c{i}", ARROW[t]) for i, t in enumerate(types)])
    data, lengths = sp.ipc_stream(schema.serialize().to_pybytes())
    sp.close()
    assert int(lengths.sum()) == len(data)
    pos = 0
    row0 = 0
    for p in range(P):
        n = int(exp["row_start"][p + 1] - exp["row_start"][p])
        if n == 0:
            assert lengths[p] == 0
            continue
        part = data[pos:pos + lengths[p]].tobytes()
        pos += int(lengths[p])
        reader = pa.ipc.open_stream(pa.py_buffer(part))
        assert reader.schema.equals(schema)
        rbs = list(reader)
        # record batch cuts: the reference's buffer policy over the rows per input batch
        counts, off = [], 0
        for b in batches:
            counts.append(int((pids[off:off + b.num_rows] == p).sum()))
            off += b.num_rows
        assert [rb.num_rows for rb in rbs] == reference_batch_rows(counts, buffer_size)
        table = pa.Table.from_batches(rbs, schema=schema)
        _check_partition([table.column(c).combine_chunks() for c in range(len(types))], exp, types, p)
    assert pos == len(data)


def _compressible_batches(types, sizes, seed):
    """Random batches whose values repeat (small ranges, a small vocabulary),
    so LZ4 blocks really compress; nulls in every nullable column."""
    from gazelle_plugin_amd.batch import HostBatch, HostColumn, bitmap_from_bools

    rng = np.random.default_rng(seed)
    words = [b"alpha", b"beta", b"gamma", b"delta-" * 3, b""]
    out = []
    for n in sizes:
        cols = []
        for t in types:
            valid = rng.random(n) > 0.1
            vb = bitmap_from_bools(valid)
            if t == T.STRING or t == T.BINARY:
                vals = [words[i] for i in rng.integers(0, len(words), n)]
                off = np.zeros(n + 1, np.int32)
                off[1:] = np.cumsum([len(v) for v in vals])
                cols.append(HostColumn(t, np.frombuffer(b"".join(vals), np.uint8).copy(), vb, off))
            elif t == T.BOOL:
                cols.append(HostColumn(t, bitmap_from_bools(rng.random(n) < 0.5), vb))
            elif t == T.NA:
                cols.append(HostColumn(t, None))
            elif t == T.DOUBLE:
                cols.append(HostColumn(t, rng.integers(0, 8, n).astype(np.float64) / 4, vb))
            else:
                cols.append(HostColumn(t, rng.integers(0, 20, n).astype(T.numpy_dtype(t)), vb))
        out.append(HostBatch(n, cols))
    return out


@pytest.mark.parametrize("compressible", [True, False])
def test_lz4_data_file_decodes_to_the_split(ctx, compressible):
    """With LZ4_FRAME bodies (SetCompressType("lz4")) every partition still
    decodes with pyarrow to the oracle's rows, batch cuts unchanged; the device
    reader decodes the same file back to the same rows."""
    from gazelle_plugin_amd.shuffle_reader import ShuffleReader

    types = [T.INT64, T.INT32, T.DOUBLE, T.STRING, T.BOOL, T.INT8, T.NA, T.BINARY]
    sizes = [70000, 0, 1, 4096, 3000]
    batches = (_compressible_batches(types, sizes, 23) if compressible
               else random_batches(types, sizes, seed=23, null_frac=0.15))
    P = 5
    exp = oracle.split(batches, types, P, [0])
    pids = oracle.split_pids(batches, types, P, [0])
    schema = pa.schema([(f"c{i}", ARROW[t]) for i, t in enumerate(types)])
    files = {}
    for codec in ("uncompressed", "lz4"):
        sp = Splitter.Make(ctx, "hash", types, P, [0], SplitOptions(buffer_size=20000))
        sp.SetCompressType(codec)
        sp.SplitBatches([DeviceBatch.from_host(b) for b in batches])
        sp.Stop()
        files[codec] = sp.ipc_stream(schema.serialize().to_pybytes())
        sp.close()
    data, lengths = files["lz4"]
    assert int(lengths.sum()) == len(data)
    if compressible:
        assert len(data) < 0.7 * len(files["uncompressed"][0])
    pos = 0
    reader = ShuffleReader(types, ctx)
    for p in range(P):
        n = int(exp["row_start"][p + 1] - exp["row_start"][p])
        part = data[pos:pos + lengths[p]]
        pos += int(lengths[p])
        if n == 0:
            assert lengths[p] == 0
            continue
        rd = pa.ipc.open_stream(pa.py_buffer(part.tobytes()))
        assert rd.schema.equals(schema)
        rbs = list(rd)
        counts, off = [], 0
        for b in batches:
            counts.append(int((pids[off:off + b.num_rows] == p).sum()))
            off += b.num_rows
        assert [rb.num_rows for rb in rbs] == reference_batch_rows(counts, 20000)
        table = pa.Table.from_batches(rbs, schema=schema)
        _check_partition([table.column(c).combine_chunks() for c in range(len(types))], exp, types, p)
        # the device reader: the decoded batches split again reproduce the partition
        dev = reader.read(part)
        assert [b.num_rows for b in dev] == [rb.num_rows for rb in rbs]
        sp = Splitter.Make(ctx, "single", types, 1)
        sp.SplitBatches(dev)
        sp.Stop()
        got = sp.read_partitions()
        sp.close()
        for c, t in enumerate(types):
            if t == T.NA:
                continue
            g = got["columns"][c]
            e = exp["columns"][c]
            r0, r1 = int(exp["row_start"][p]), int(exp["row_start"][p + 1])
            gv = np.ones(n, bool) if g["validity"] is None else _bits(g["validity"], 0, n)
            ev = np.ones(n, bool) if e.get("validity") is None else _bits(e["validity"], r0, n)
            np.testing.assert_array_equal(gv, ev, err_msg=f"validity p={p} c={c}")
    assert pos == len(data)


def test_lz4_encoder_block_format_cases(ctx):
    """The device encoder on the block format's cases (incompressible data, runs
    of periods 1..70000 with matches longer than one sequence takes, long
    literal runs, inputs below the 13-byte minimum, several 64 KiB blocks):
    each case is the one value of a single-partition binary column, so its
    bytes are one body buffer; pyarrow's LZ4 decoder must give them back."""
    from test_oracle_lz4 import _cases

    cases = [d for _, d in _cases()]
    types = [T.BINARY]
    schema = pa.schema([("v", pa.binary())])
    from gazelle_plugin_amd.batch import HostBatch, HostColumn

    for d in cases:
        off = np.array([0, len(d)], np.int32)
        hb = HostBatch(1, [HostColumn(T.BINARY, np.frombuffer(d, np.uint8).copy(), None, off)])
        # one-row batches: threshold 0 so the encoder runs (the default leaves
        # batches of at most 256 rows uncompressed)
        sp = Splitter.Make(ctx, "single", types, 1, [], SplitOptions(batch_compress_threshold=0))
        sp.SetCompressType("lz4")
        sp.Split(DeviceBatch.from_host(hb))
        sp.Stop()
        data, lengths = sp.ipc_stream(schema.serialize().to_pybytes())
        sp.close()
        rbs = list(pa.ipc.open_stream(pa.py_buffer(data.tobytes())))
        assert len(rbs) == 1 and rbs[0].column(0).to_pylist() == [d]
        if len(d) >= 1 << 20:  # a megabyte of zeros: 16 blocks that really compress
            assert len(data) < len(d) // 50
