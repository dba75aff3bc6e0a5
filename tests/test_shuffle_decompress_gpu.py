"""LZ4_FRAME shuffle streams decoded on the GPU: every buffer of every record
batch of a pyarrow-compressed stream (the lz4 library the reference's Arrow
uses) equals, byte for byte, the oracle's decode of the same compressed buffer;
the decoded batches split again on the device reproduce the oracle's split of
pyarrow's decoded rows; corrupt data and a wrong length prefix give the
reference's error statuses."""
import numpy as np
import pyarrow as pa
import pytest

import oracle
from gazelle_plugin_amd import _lib as L
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.batch import bitmap_from_bools, bools_from_bitmap, from_arrow
from gazelle_plugin_amd.hip import to_host
from gazelle_plugin_amd.shuffle_reader import ShuffleReader
from gazelle_plugin_amd.splitter import Splitter
from splitcheck import assert_layout_equal
from test_ipc_reader import TYPES, _batches, _stream

pytestmark = pytest.mark.gpu


def _buffer_spans(metadata: bytes):
    """(offset, length) of each body buffer, read from the RecordBatch message
    flatbuffer (Message.header -> RecordBatch.buffers)."""
    b = bytes(metadata)

    def u(p, n, signed=False):
        return int.from_bytes(b[p:p + n], "little", signed=signed)

    def field(t, i):
        vt = t - u(t, 4, True)
        if 4 + 2 * i + 2 > u(vt, 2):
            return None
        fo = u(vt + 4 + 2 * i, 2)
        return t + fo if fo else None

    root = u(0, 4)
    hdr = field(root, 2)
    rb = hdr + u(hdr, 4)
    bf = field(rb, 2)
    vec = bf + u(bf, 4)
    return [(u(vec + 4 + 16 * i, 8, True), u(vec + 12 + 16 * i, 8, True)) for i in range(u(vec, 4))]


def _oracle_buffers(data, types):
    """Per record batch, the JNI-order buffers decoded by the oracle (None for
    empty ones)."""
    out = []
    for msg in pa.ipc.MessageReader.open_stream(pa.py_buffer(data)):
        if msg.type != "record batch":
            continue
        body = np.frombuffer(msg.body, np.uint8)
        spans = iter(_buffer_spans(msg.metadata))
        bufs = []
        for t in types:
            if t == T.NA:
                bufs += [None, None]
                continue
            for _ in range(3 if T.is_binary(t) else 2):
                off, n = next(spans)
                bufs.append(oracle.ipc_decompress_buffer(body[off:off + n]) if n else None)
        out.append(bufs)
    return out


def _assert_decoded(batches, expected):
    assert len(batches) == len(expected)
    for b, (db, exp) in enumerate(zip(batches, expected)):
        for k, e in enumerate(exp):
            if e is None or len(e) == 0:
                assert db.sizes[k] == 0, f"batch {b} buffer {k}"
                continue
            assert db.sizes[k] == len(e), f"batch {b} buffer {k}"
            np.testing.assert_array_equal(to_host(db.addrs[k], db.sizes[k]), e, err_msg=f"batch {b} buffer {k}")


@pytest.mark.parametrize("null_frac", [0.0, 0.2])
def test_lz4_stream_decodes_like_the_oracle(ctx, null_frac):
    schema, rbs = _batches(TYPES, [100_000, 1, 4096, 777], seed=11, null_frac=null_frac)
    data = _stream(rbs, schema, compression="lz4")
    batches = ShuffleReader(TYPES, ctx).read(data)
    assert [b.num_rows for b in batches] == [rb.num_rows for rb in rbs]
    _assert_decoded(batches, _oracle_buffers(data, TYPES))


def test_lz4_multiblock_buffers(ctx):
    """Buffers of 800 KB are 13 linked 64 KiB blocks whose matches cross blocks;
    runs and periodic data give overlapping matches."""
    types = [T.INT64, T.INT32, T.STRING]
    schema = pa.schema([("a", pa.int64()), ("b", pa.int32()), ("c", pa.string())])
    n = 100_000
    rb = pa.record_batch([pa.array(np.repeat(np.arange(n // 5, dtype=np.int64), 5)),
                          pa.array(np.zeros(n, np.int32)),
                          pa.array(["abc" * (i % 7) for i in range(n)])], schema=schema)
    data = _stream([rb], schema, compression="lz4")
    batches = ShuffleReader(types, ctx).read(data)
    _assert_decoded(batches, _oracle_buffers(data, types))


@pytest.mark.parametrize("batched", [True, False])
def test_decoded_batches_split_like_the_oracle(ctx, batched):
    schema, rbs = _batches(TYPES, [5000, 3, 4096], seed=12, null_frac=0.15)
    data = _stream(rbs, schema, compression="lz4")
    dev = ShuffleReader(TYPES, ctx).read(data)
    sp = Splitter.Make(ctx, "single", TYPES, 1)
    if batched:
        sp.SplitBatches(dev)
    else:
        for b in dev:
            sp.Split(b)
    sp.Stop()
    got = sp.read_partitions()
    sp.close()
    host = [_raw_host(rb) for rb in pa.ipc.open_stream(pa.py_buffer(data))]
    assert_layout_equal(got, oracle.split(host, TYPES, 1, [], mode="single"), TYPES)


def _raw_host(rb):
    """from_arrow, with boolean value bits kept as stored under nulls too (the
    split moves them unchanged)."""
    hb = from_arrow(rb)
    for c, col in enumerate(hb.columns):
        if col.type_id == T.BOOL:
            raw = np.frombuffer(rb.column(c).buffers()[1], np.uint8)
            col.values = bitmap_from_bools(bools_from_bitmap(raw, rb.num_rows))
    return hb


def _frame_prefix(data):
    """Stream offset of the length prefix in front of the first LZ4 frame."""
    pos = data.tobytes().find(b"\x04\x22\x4d\x18")
    assert pos >= 8
    return pos - 8


def test_wrong_length_prefix_is_invalid(ctx):
    schema, rbs = _batches([T.INT64], [5000], seed=13, null_frac=0)
    data = _stream(rbs, schema, compression="lz4").copy()
    p = _frame_prefix(data)
    n = int(data[p:p + 8].view(np.int64)[0])
    data[p:p + 8] = np.frombuffer(np.int64(n + 8).tobytes(), np.uint8)
    with pytest.raises(L.Invalid):
        ShuffleReader([T.INT64], ctx).read(data)


def test_corrupt_block_is_io_error(ctx):
    schema = pa.schema([("a", pa.int64())])
    rb = pa.record_batch([pa.array(np.arange(5000, dtype=np.int64))], schema=schema)
    data = _stream([rb], schema, compression="lz4").copy()
    p = _frame_prefix(data)
    assert data[p + 12] & 0x08 == 0  # no content size in the frame header
    blk = p + 8 + 7 + 4  # prefix, frame header, block size word
    assert data[blk - 1] & 0x80 == 0  # a compressed (not stored) block
    data[blk:blk + 64] = 0xFF  # a literal length that runs past its block
    body = next(m for m in pa.ipc.MessageReader.open_stream(pa.py_buffer(data)) if m.type == "record batch").body
    off, n = _buffer_spans(next(m for m in pa.ipc.MessageReader.open_stream(pa.py_buffer(data))
                                if m.type == "record batch").metadata)[1]
    with pytest.raises(ValueError):  # the oracle rejects it too
        oracle.ipc_decompress_buffer(np.frombuffer(body, np.uint8)[off:off + n])
    with pytest.raises(L.GazelleError) as e:
        ShuffleReader([T.INT64], ctx).read(data)
    assert e.value.code == 5


def test_block_format_cases(ctx):
    """The oracle's LZ4 cases (stored blocks, runs with periods 1..70000, long
    length extensions, linked blocks) as binary values, one record batch each."""
    from test_oracle_lz4 import _cases

    cases = [d for _, d in _cases()]
    schema = pa.schema([("v", pa.binary())])
    rbs = [pa.record_batch([pa.array([d], type=pa.binary())], schema=schema) for d in cases]
    data = _stream(rbs, schema, compression="lz4")
    batches = ShuffleReader([T.BINARY], ctx).read(data)
    _assert_decoded(batches, _oracle_buffers(data, [T.BINARY]))
    for db, d in zip(batches, cases):
        got = to_host(db.addrs[2], db.sizes[2]) if db.sizes[2] else np.zeros(0, np.uint8)
        assert got.tobytes() == d
