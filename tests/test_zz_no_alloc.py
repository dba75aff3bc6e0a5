"""The allocation-free LZ4_FRAME decode (gz_ipc_decompress_noalloc) against the
allocating one (gz_ipc_decompress), through the C ABI: byte-equal arenas and
equal statuses on two calls in a row, the second with a smaller stream, and
the reference's error status for a corrupt block."""
import ctypes

import numpy as np
import pytest
import torch

from gazelle_plugin_amd import _lib as L
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.shuffle_reader import parse_stream
from test_ipc_reader import TYPES, _batches, _stream

pytestmark = pytest.mark.gpu


def _decode(ctx, fn, data, types):
    dev = torch.from_numpy(data).to("cuda")
    base = int(dev.data_ptr())
    _, _, _, arena_bytes = parse_stream(data, base, types)
    arena = torch.full((max(1, arena_bytes),), 0xA5, dtype=torch.uint8, device="cuda")
    rc = fn(ctx.handle, data.ctypes.data_as(ctypes.c_void_p), len(data), ctypes.c_int64(base), len(types),
            L.i32_array(types), ctypes.c_void_p(arena.data_ptr()), ctypes.c_int64(arena_bytes))
    ctx.synchronize()
    return rc, arena.cpu().numpy()[:arena_bytes], arena_bytes


def test_noalloc_decode_equals_allocating_decode(ctx):
    lib = L.lib()
    schema, rbs = _batches(TYPES, [60_000, 1, 4096, 777], seed=21, null_frac=0.2)
    big = _stream(rbs, schema, compression="lz4")
    schema2, rbs2 = _batches(TYPES, [3000, 5], seed=22, null_frac=0.1)
    small = _stream(rbs2, schema2, compression="lz4")
    assert len(small) < len(big)
    for data in (big, small):  # two calls in a row, the second smaller
        rc_a, arena_a, n_a = _decode(ctx, lib.gz_ipc_decompress, data, TYPES)
        rc_b, arena_b, n_b = _decode(ctx, lib.gz_ipc_decompress_noalloc, data, TYPES)
        assert rc_a == rc_b == 0
        assert n_a == n_b > 0
        assert np.array_equal(arena_a, arena_b)


def test_noalloc_decode_reports_corrupt_block(ctx):
    import pyarrow as pa

    lib = L.lib()
    schema = pa.schema([("a", pa.int64())])
    rb = pa.record_batch([pa.array(np.arange(5000, dtype=np.int64))], schema=schema)
    data = _stream([rb], schema, compression="lz4").copy()
    p = data.tobytes().find(b"\x04\x22\x4d\x18") - 8
    data[p + 8 + 7 + 4:p + 8 + 7 + 4 + 64] = 0xFF  # a literal length that runs past its block
    rc_a, _, _ = _decode(ctx, lib.gz_ipc_decompress, data, [T.INT64])
    rc_b, _, _ = _decode(ctx, lib.gz_ipc_decompress_noalloc, data, [T.INT64])
    assert rc_a == rc_b == 5  # IOError
