"""The oracle's LZ4 frame decoder pinned by pyarrow's LZ4_FRAME codec (the lz4
library the reference's Arrow uses): frames pyarrow compresses decode to the
original bytes across the block format's cases -- literals only, stored
(incompressible) blocks, overlapping matches (run lengths with offsets 1..7),
long length extensions, linked 64 KiB blocks whose matches cross block
boundaries -- and truncated or corrupted frames are rejected, never overrun."""
import numpy as np
import pyarrow as pa
import pytest

import oracle


def _lz4(b: bytes) -> bytes:
    return pa.Codec("lz4").compress(b, asbytes=True)


def _cases():
    rng = np.random.default_rng(5)
    yield "empty", b""
    for n in (1, 5, 12, 13, 15, 16, 19, 20, 64, 65):
        yield f"random{n}", rng.integers(0, 256, n, dtype=np.uint8).tobytes()
    yield "random_200k", rng.integers(0, 256, 200_000, dtype=np.uint8).tobytes()  # stored blocks
    yield "zeros_1M", bytes(1 << 20)  # offset-1 runs, long extensions
    for period in (2, 3, 5, 7, 8, 13, 300, 70000):
        pat = rng.integers(0, 256, period, dtype=np.uint8)
        yield f"period{period}", np.resize(pat, 150_001).tobytes()
    yield "arange_i64", np.arange(100_000, dtype=np.int64).tobytes()
    yield "offsets_i32", np.cumsum(rng.integers(0, 9, 65_537)).astype(np.int32).tobytes()
    words = [b"alpha", b"beta", b"gamma", b"delta", b"x" * 40, b""]
    yield "text", b"".join(words[i] for i in rng.integers(0, len(words), 50_000))
    yield "block_edge", rng.integers(0, 4, 65536 * 3 + 1, dtype=np.uint8).tobytes()


@pytest.mark.parametrize("name,data", list(_cases()), ids=[c[0] for c in _cases()])
def test_decodes_pyarrow_frames(name, data):
    frame = _lz4(data)
    got = oracle.lz4f_decompress(frame, len(data))
    assert got.tobytes() == data


def test_concatenated_frames():
    a, b = b"hello " * 1000, bytes(range(256)) * 300
    got = oracle.lz4f_decompress(_lz4(a) + _lz4(b), len(a) + len(b))
    assert got.tobytes() == a + b


def test_rejects_truncated_and_corrupted():
    data = np.arange(50_000, dtype=np.int64).tobytes()
    frame = _lz4(data)
    for cut in (3, 6, 7, 11, len(frame) // 2, len(frame) - 1):
        with pytest.raises(ValueError):
            oracle.lz4f_decompress(frame[:cut], len(data))
    with pytest.raises(ValueError):  # output capacity too small
        oracle.lz4f_decompress(frame, len(data) - 1)
    bad = bytearray(frame)
    bad[0] ^= 1  # magic
    with pytest.raises(ValueError):
        oracle.lz4f_decompress(bytes(bad), len(data))


def test_ipc_buffer_prefix():
    data = np.arange(1000, dtype=np.int32).tobytes()
    buf = np.int64(len(data)).tobytes() + _lz4(data)
    assert oracle.ipc_decompress_buffer(buf).tobytes() == data
    raw = np.int64(-1).tobytes() + data
    assert oracle.ipc_decompress_buffer(raw).tobytes() == data
