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


# ---- Spark hashUnsafeBytes (utf8 / binary keys, hashbuf_spark)
#
# Spark's Murmur3_x86_32.hashUnsafeBytes mixes the whole 4-byte words as
# Murmur3 does, then every tail byte on its own as one more block holding the
# byte sign-extended to an int (Platform.getByte), and finishes with
# fmix(h, length).  Murmur3 over the same words with each tail byte expanded to
# that 4-byte block differs only in the length fmix xors in; fmix's mixing part
# is invertible, so the expected value follows from scikit-learn's independent
# Murmur3 over the expanded bytes: undo its finalizer, swap the length, redo it.

M32 = 0xFFFFFFFF


def _fmix_core(h):
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & M32
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & M32
    return h ^ (h >> 16)


def _fmix_core_inverse(h):
    h ^= h >> 16
    h = (h * pow(0xC2B2AE35, -1, 1 << 32)) & M32
    h ^= (h >> 13) ^ (h >> 26)
    h = (h * pow(0x85EBCA6B, -1, 1 << 32)) & M32
    return h ^ (h >> 16)


def spark_hash_unsafe_bytes(b: bytes, seed: int) -> int:
    q = len(b) - len(b) % 4
    expanded = b[:q] + b"".join(int(np.int8(np.uint8(x))).to_bytes(4, "little", signed=True) for x in b[q:])
    h = _fmix_core_inverse(mm3(expanded, seed & M32) & M32) ^ len(expanded)  # Murmur3's state before fmix
    return int(np.int32(np.uint32(_fmix_core(h ^ len(b)))))


def test_fmix_inverse_is_an_inverse():
    rng = np.random.default_rng(11)
    for x in [0, 1, M32, 0x80000000] + [int(v) for v in rng.integers(0, 1 << 32, 2000, dtype=np.uint64)]:
        assert _fmix_core(_fmix_core_inverse(x)) == x and _fmix_core_inverse(_fmix_core(x)) == x


@pytest.mark.parametrize("seed", [42, 0, -7])
def test_hash_bytes_tail_is_spark_hash_unsafe_bytes(seed):
    """Every length 0..40 (every tail length), bytes >= 0x80 in the tail (the
    sign extension), and chained seeds as a multi-key hash uses them."""
    rng = np.random.default_rng(12 + (seed & 0xFF))
    for n in range(41):
        for _ in range(25):
            b = bytes(rng.integers(0, 256, size=n, dtype=np.uint8))
            assert oracle.hash_bytes(b, seed) == spark_hash_unsafe_bytes(b, seed), (n, b.hex())
    for b in (b"\x80", b"\xff\xff\xff", b"abcd\x80", b"a", b"ab", b"abc", "é".encode()):
        assert oracle.hash_bytes(b, seed) == spark_hash_unsafe_bytes(b, seed)


def test_tail_rule_differs_from_standard_murmur3():
    """The pin is not vacuous: standard Murmur3 (zero-extended, one tail block)
    disagrees with Spark's rule on tails, and so would a zero-extension."""
    b = b"abcde\xf0"
    assert spark_hash_unsafe_bytes(b, 42) != mm3(b, 42)
    zero_ext = b[:4] + b"".join(x.to_bytes(4, "little") for x in b[4:])
    h = _fmix_core_inverse(mm3(zero_ext, 42) & M32) ^ len(zero_ext)
    assert spark_hash_unsafe_bytes(b, 42) != int(np.int32(np.uint32(_fmix_core(h ^ len(b)))))


# ---- float / double keys (hash32_spark / hash64_spark on floatToIntBits / doubleToLongBits)

def test_float_and_double_keys_hash_their_bits():
    """An ordinary float or double key hashes as Murmur3 over its little-endian
    bits; -0.0 and NaN (version-dependent in Spark) are refused."""
    from gazelle_plugin_amd import types as T
    from gazelle_plugin_amd.batch import HostBatch, HostColumn

    rng = np.random.default_rng(13)
    d = np.concatenate([rng.standard_normal(500) * 1e6, [0.0, 1.0, -1.0, np.inf, -np.inf, 5e-324]])
    b = HostBatch(len(d), [HostColumn(T.DOUBLE, d)])
    assert list(oracle.split_pids([b], [T.DOUBLE], 200, [0])) == [mm3(x.tobytes(), 42) % 200 for x in d]
    f = d.astype(np.float32)
    b = HostBatch(len(f), [HostColumn(T.FLOAT, f)])
    assert list(oracle.split_pids([b], [T.FLOAT], 200, [0])) == [mm3(x.tobytes(), 42) % 200 for x in f]
    for bad in (-0.0, np.nan):
        for t, dt in ((T.DOUBLE, np.float64), (T.FLOAT, np.float32)):
            b = HostBatch(2, [HostColumn(t, np.array([1.0, bad], dt))])
            with pytest.raises(ValueError):
                oracle.split_pids([b], [t], 200, [0])
