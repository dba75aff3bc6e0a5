"""The shuffle read's stream parser against pyarrow's IPC reader, an independent
implementation of the format: for streams pyarrow writes, every buffer address
the parser gives (taking the stream's own host address as the device base) is
the address of the buffer pyarrow's zero-copy reader returns, with its length;
batch row counts match; schema messages, the legacy framing and empty streams
are handled; compressed bodies, truncation and a wrong field count are errors
with the reference's status codes."""
import numpy as np
import pyarrow as pa
import pytest

from gazelle_plugin_amd import _lib as L
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.shuffle_reader import parse_stream

TYPES = [T.INT64, T.STRING, T.BOOL, T.DOUBLE, T.NA, T.INT8, T.BINARY, T.DATE32, T.UINT16]
ARROW = {T.INT64: pa.int64(), T.STRING: pa.string(), T.BOOL: pa.bool_(), T.DOUBLE: pa.float64(),
         T.NA: pa.null(), T.INT8: pa.int8(), T.BINARY: pa.binary(), T.DATE32: pa.date32(),
         T.UINT16: pa.uint16(), T.INT32: pa.int32()}


def _column(t, n, rng, null_frac):
    mask = rng.random(n) < null_frac if null_frac else None
    if t == T.NA:
        return pa.nulls(n)
    if t in (T.STRING, T.BINARY):
        vals = ["x" * int(k) for k in rng.integers(0, 9, n)]
        a = pa.array(vals, type=pa.string(), mask=mask)
        return a if t == T.STRING else a.cast(pa.binary())
    if t == T.BOOL:
        return pa.array(rng.random(n) < 0.5, mask=mask)
    if t == T.DOUBLE:
        return pa.array(rng.standard_normal(n), mask=mask)
    if t == T.DATE32:
        return pa.array(rng.integers(0, 20000, n).astype(np.int32), type=pa.date32(), mask=mask)
    dt = T.numpy_dtype(t)
    info = np.iinfo(dt)
    return pa.array(rng.integers(info.min, info.max, n, dtype=dt, endpoint=True), mask=mask)


def _stream(batches, schema, **opts):
    sink = pa.BufferOutputStream()
    with pa.ipc.new_stream(sink, schema, options=pa.ipc.IpcWriteOptions(**opts)) as w:
        for rb in batches:
            w.write_batch(rb)
    return np.frombuffer(sink.getvalue().to_pybytes(), np.uint8).copy()


def _batches(types, sizes, seed, null_frac):
    rng = np.random.default_rng(seed)
    schema = pa.schema([(f"c{i}", ARROW[t]) for i, t in enumerate(types)])
    return schema, [pa.record_batch([_column(t, n, rng, null_frac) for t in types], schema=schema)
                    for n in sizes]


def _expected(rb, types):
    """(address, length) per JNI-order buffer of pyarrow's zero-copy batch."""
    out = []
    for a, t in zip(rb.columns, types):
        bufs = a.buffers()
        if t == T.NA:
            out += [(0, 0), (0, 0)]
            continue
        for b in bufs:
            out.append((0, 0) if b is None or b.size == 0 else (b.address, b.size))
    return out


@pytest.mark.parametrize("legacy", [False, True])
@pytest.mark.parametrize("null_frac", [0.0, 0.2])
def test_addresses_match_pyarrow(legacy, null_frac):
    schema, rbs = _batches(TYPES, [1000, 1, 0, 4096, 77], seed=3, null_frac=null_frac)
    data = _stream(rbs, schema, use_legacy_format=legacy)
    base = data.ctypes.data
    rows, addrs, sizes, _ = parse_stream(data, base, TYPES)
    read = list(pa.ipc.open_stream(pa.py_buffer(data)))
    assert rows.tolist() == [rb.num_rows for rb in rbs] == [rb.num_rows for rb in read]
    for b, rb in enumerate(read):
        exp = _expected(rb, TYPES)
        got = list(zip(addrs[b].tolist(), sizes[b].tolist()))
        assert got == exp, f"batch {b}"
        assert all(a % 8 == 0 for a, _ in got), "buffers stay 8-byte aligned"


def test_offsets_without_base():
    schema, rbs = _batches([T.INT32, T.STRING], [10], seed=4, null_frac=0.5)
    data = _stream(rbs, schema)
    rows, addrs, sizes, _ = parse_stream(data, 0, [T.INT32, T.STRING])
    rb = next(iter(pa.ipc.open_stream(pa.py_buffer(data))))
    vals = rb.column(0).buffers()[1]
    k = 1  # values buffer of column 0
    np.testing.assert_array_equal(data[addrs[0, k]:addrs[0, k] + sizes[0, k]],
                                  np.frombuffer(vals, np.uint8)[: sizes[0, k]])


def test_schema_only_stream_has_no_batches():
    schema, _ = _batches([T.INT64], [], seed=5, null_frac=0)
    rows, addrs, sizes, _ = parse_stream(_stream([], schema), 0, [T.INT64])
    assert len(rows) == 0


def test_stream_without_eos_ends_at_its_length():
    schema, rbs = _batches([T.INT64], [5, 6], seed=6, null_frac=0)
    data = _stream(rbs, schema)[:-8]  # drop the end-of-stream marker
    rows, _, _, _ = parse_stream(data, 0, [T.INT64])
    assert rows.tolist() == [5, 6]


def test_zstd_stream_is_not_implemented():
    schema, rbs = _batches([T.INT64, T.STRING], [100], seed=7, null_frac=0.1)
    data = _stream(rbs, schema, compression="zstd")
    with pytest.raises(L.NotImplementedStatus):
        parse_stream(data, 0, [T.INT64, T.STRING])


def test_lz4_stream_lays_out_the_arena():
    """Compressed buffers go to 64-byte aligned arena slots of their decoded
    size (pyarrow's decode gives the sizes); empty buffers stay empty."""
    schema, rbs = _batches(TYPES, [1000, 1, 4096], seed=10, null_frac=0.2)
    data = _stream(rbs, schema, compression="lz4")
    rows, addrs, sizes, arena = parse_stream(data, 1 << 40, TYPES, arena_base=1 << 44)
    read = list(pa.ipc.open_stream(pa.py_buffer(data)))
    assert rows.tolist() == [rb.num_rows for rb in read]
    slots = []
    for b, rb in enumerate(read):
        for (ea, es), a, s in zip(_expected(rb, TYPES), addrs[b], sizes[b]):
            if es == 0:
                assert a == 0 and s == 0
                continue
            assert a >= 1 << 44 and (a - (1 << 44)) % 64 == 0
            assert s >= es  # pyarrow may trim a decoded bitmap or offsets view; the stored length is the prefix
            slots.append((a - (1 << 44), s))
    slots.sort()
    for (a0, s0), (a1, _) in zip(slots, slots[1:]):
        assert a0 + s0 <= a1
    assert arena >= slots[-1][0] + slots[-1][1]


def test_truncated_stream_is_io_error():
    schema, rbs = _batches([T.INT64, T.STRING], [100], seed=8, null_frac=0.1)
    data = _stream(rbs, schema)
    for cut in (len(data) - 9, len(data) - 100, 20):
        with pytest.raises(L.GazelleError) as e:
            parse_stream(data[:cut], 0, [T.INT64, T.STRING])
        assert e.value.code == 5  # IOError


def test_field_count_mismatch_is_invalid():
    schema, rbs = _batches([T.INT64, T.STRING], [10], seed=9, null_frac=0)
    with pytest.raises(L.Invalid):
        parse_stream(_stream(rbs, schema), 0, [T.INT64])
