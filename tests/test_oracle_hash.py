"""Pins of the oracle's Spark Murmur3 hashing and pmod.

Spark's hashInt / hashLong are MurmurHash3_x86_32 over the 4 / 8 little-endian
bytes of the value; scikit-learn's murmurhash3_32 is an independent
implementation of that published algorithm, so agreement on random keys pins the
oracle (a dropped mixing step, a swapped word order or a wrong constant fails).
The generic function is itself pinned to the published test vectors of
MurmurHash3_x86_32.
"""
import numpy as np
import pytest

import oracle

sk = pytest.importorskip("sklearn.utils")


def mm3(b: bytes, seed: int) -> int:
    return int(sk.murmurhash3_32(b, seed=seed, positive=False))


def test_reference_library_known_answers():
    # published MurmurHash3_x86_32 vectors (as unsigned 32-bit)
    u = lambda b, s: mm3(b, s) & 0xFFFFFFFF
    assert u(b"", 0) == 0
    assert u(b"", 1) == 0x514E28B7
    assert u(b"", 0xFFFFFFFF) == 0x81F16F39
    assert u(b"\x00\x00\x00\x00", 0) == 0x2362F9DE
    assert u(b"Hello, world!", 0x9747B28C) == 0x24884CBA
    assert u(b"The quick brown fox jumps over the lazy dog", 0x9747B28C) == 0x2FA826CD


@pytest.mark.parametrize("seed", [42, 0, -1, 12345])
def test_hash_long_is_murmur3_of_le_bytes(seed):
    rng = np.random.default_rng(7)
    keys = list(rng.integers(-(1 << 63), (1 << 63) - 1, size=3000, dtype=np.int64))
    keys += [0, 1, -1, (1 << 63) - 1, -(1 << 63), 1 << 40, (1 << 32) - 1, 1 << 32]
    s32 = int(np.int32(np.uint32(seed & 0xFFFFFFFF)))  # the seed as Java's int
    for k in keys:
        assert oracle.hash_long(int(k), s32) == mm3(np.int64(k).tobytes(), seed & 0xFFFFFFFF), k


@pytest.mark.parametrize("seed", [42, 0, 7])
def test_hash_int_is_murmur3_of_le_bytes(seed):
    rng = np.random.default_rng(8)
    for k in list(rng.integers(-(1 << 31), (1 << 31) - 1, size=3000)) + [0, -1, 2**31 - 1, -2**31]:
        assert oracle.hash_int(int(k), seed) == mm3(np.int32(k).tobytes(), seed)


def test_hash_bytes_whole_words_are_murmur3():
    # for lengths that are multiples of 4 Spark's hashUnsafeBytes is plain Murmur3
    rng = np.random.default_rng(9)
    for n in (0, 4, 8, 12, 32):
        for _ in range(50):
            b = bytes(rng.integers(0, 256, size=n, dtype=np.uint8))
            assert oracle.hash_bytes(b, 42) == mm3(b, 42)


def test_pmod_matches_floored_modulo():
    rng = np.random.default_rng(10)
    hs = list(rng.integers(-(1 << 31), (1 << 31) - 1, size=5000)) + [-2**31, 2**31 - 1, 0, -1]
    for P in (1, 2, 3, 200, 4096, 65536):
        for h in hs:
            assert oracle.pmod(int(h), P) == int(h) % P


def test_split_pids_follow_hash_and_pmod():
    from gazelle_plugin_amd import types as T
    from gazelle_plugin_amd.batch import HostBatch, HostColumn

    keys = np.array([0, 1, -1, 1 << 40, 123456789012], np.int64)
    b = HostBatch(len(keys), [HostColumn(T.INT64, keys)])
    pids = oracle.split_pids([b], [T.INT64], 200, [0])
    exp = [(mm3(np.int64(k).tobytes(), 42)) % 200 for k in keys]
    assert list(pids) == exp


def test_null_key_keeps_seed():
    from gazelle_plugin_amd import types as T
    from gazelle_plugin_amd.batch import HostBatch, HostColumn, bitmap_from_bools

    keys = np.array([5, 7], np.int64)
    b = HostBatch(2, [HostColumn(T.INT64, keys, bitmap_from_bools(np.array([False, True])))])
    pids = oracle.split_pids([b], [T.INT64], 200, [0])
    assert pids[0] == 42 % 200
    assert pids[1] == mm3(np.int64(7).tobytes(), 42) % 200
