"""The partition writer's record-batch metadata (a hand-built Arrow Message
flatbuffer) read back by pyarrow's IPC reader, an independent implementation of
the format: a stream of the schema, one record batch (int64 with nulls, utf8,
boolean, float64 without nulls) and the end-of-stream marker decodes to the
arrays the body holds."""
import numpy as np
import pyarrow as pa
import pytest

from gazelle_plugin_amd.partition_writer import record_batch_message


def _pad8(b: bytes) -> bytes:
    return b + b"\0" * (-len(b) % 8)


def _body(arrays):
    """Arrow IPC body of the arrays (validity only with nulls, buffers padded to
    8) and its field nodes and buffer table."""
    body, nodes, bufs = b"", [], []

    def add(b):
        nonlocal body
        bufs.append((len(body), len(b)))
        body += _pad8(b)

    for a in arrays:
        nodes.append((len(a), a.null_count))
        vb = a.buffers()[0]
        if a.null_count > 0:
            add(vb.to_pybytes()[: (len(a) + 7) // 8])
        else:
            bufs.append((len(body), 0))
        for b in a.buffers()[1:]:
            if pa.types.is_string(a.type) and b is a.buffers()[1]:
                add(b.to_pybytes()[: 4 * (len(a) + 1)])
            elif pa.types.is_string(a.type):
                n = int(np.frombuffer(a.buffers()[1], np.int32)[len(a)])
                add(b.to_pybytes()[:n])
            elif pa.types.is_boolean(a.type):
                add(b.to_pybytes()[: (len(a) + 7) // 8])
            else:
                add(b.to_pybytes()[: len(a) * a.type.bit_width // 8])
    return body, nodes, bufs


@pytest.mark.parametrize("n", [0, 1, 7, 8, 1000])
def test_record_batch_message_decodes(n):
    rng = np.random.default_rng(n)
    i64 = pa.array([None if rng.random() < 0.2 else int(x) for x in rng.integers(-9, 9, n)], pa.int64())
    s = pa.array(["x" * int(k) for k in rng.integers(0, 12, n)], pa.string())
    b = pa.array([bool(x) for x in rng.integers(0, 2, n)], pa.bool_())
    f = pa.array(rng.normal(size=n), pa.float64())
    schema = pa.schema([("k", i64.type), ("s", s.type), ("b", b.type), ("f", f.type)])
    body, nodes, bufs = _body([i64, s, b, f])
    msg = record_batch_message(n, nodes, bufs, len(body))
    assert len(msg) % 8 == 0
    stream = schema.serialize().to_pybytes() + msg + body + b"\xff\xff\xff\xff\0\0\0\0"
    got = pa.ipc.open_stream(pa.py_buffer(stream)).read_all()
    exp = pa.Table.from_arrays([i64, s, b, f], schema=schema)
    assert got.equals(exp)
