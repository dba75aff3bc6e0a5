"""Seeded synthetic inputs of the benchmark workloads.

Seeds are numpy PCG64 with seed = 42 + column index (as the measurement plan
states); batches are 4096 rows.  ``*_host`` functions return HostBatch lists
(numpy) for parity tests; ``*_device`` functions build the full-size inputs
directly in HBM with torch and return DeviceBatch views into them.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

from . import types as T
from .batch import DeviceBatch, HostBatch, HostColumn, bitmap_from_bools

BATCH_ROWS = 4096

SPLIT_TYPES = [T.INT64, T.INT32, T.DOUBLE, T.STRING]
AGG_TYPES = [T.INT64, T.DOUBLE, T.INT32]


def _rng(i: int, seed: int = 42):
    return np.random.Generator(np.random.PCG64(seed + i))


# ------------------------------------------------------------------ split workload

def split_host(num_rows: int, batch_rows: int = BATCH_ROWS, seed: int = 42,
               null_frac: float = 0.02) -> List[HostBatch]:
    """int64 key U[0,2^40), int32 uniform, float64 N(0,1) with 2% nulls,
    utf8 of length U[8,32] over [a-z]."""
    key = _rng(0, seed).integers(0, 1 << 40, size=num_rows, dtype=np.int64)
    i32 = _rng(1, seed).integers(-(1 << 31), (1 << 31) - 1, size=num_rows, dtype=np.int64).astype(np.int32)
    r2 = _rng(2, seed)
    f64 = r2.standard_normal(num_rows)
    f64_valid = r2.random(num_rows) >= null_frac
    r3 = _rng(3, seed)
    lens = r3.integers(8, 33, size=num_rows).astype(np.int32)
    chars = r3.integers(97, 123, size=int(lens.sum()), dtype=np.uint8)
    starts = np.zeros(num_rows + 1, np.int64)
    starts[1:] = np.cumsum(lens)
    out = []
    for b0 in range(0, num_rows, batch_rows):
        b1 = min(num_rows, b0 + batch_rows)
        n = b1 - b0
        off = (starts[b0:b1 + 1] - starts[b0]).astype(np.int32)
        data = chars[starts[b0]:starts[b1]].copy()
        valid = f64_valid[b0:b1]
        out.append(HostBatch(n, [
            HostColumn(T.INT64, key[b0:b1].copy()),
            HostColumn(T.INT32, i32[b0:b1].copy()),
            HostColumn(T.DOUBLE, f64[b0:b1].copy(), bitmap_from_bools(valid) if not valid.all() else None),
            HostColumn(T.STRING, data, None, off),
        ]))
    return out


class DeviceTable:
    """Whole-workload device tensors plus DeviceBatch views of every batch."""

    def __init__(self, tensors, batches: List[DeviceBatch], num_rows: int, nbytes: int):
        self.tensors = tensors
        self.batches = batches
        self.num_rows = num_rows
        self.nbytes = nbytes


def _bitmap_device(valid_bool):
    """torch bool [n] (n multiple of 8) -> uint8 LSB-first bitmap on the device."""
    import torch

    n = valid_bool.numel()
    v = valid_bool.view(-1, 8).to(torch.uint8)
    w = torch.tensor([1, 2, 4, 8, 16, 32, 64, 128], dtype=torch.uint8, device=v.device)
    return (v.to(torch.int32) * w.to(torch.int32)).sum(dim=1).to(torch.uint8)


def split_device(num_rows: int, batch_rows: int = BATCH_ROWS, seed: int = 42,
                 device: str = "cuda") -> DeviceTable:
    """The shuffle-split workload generated in HBM (torch Philox, seeded)."""
    import torch

    assert num_rows % batch_rows == 0 and batch_rows % 64 == 0
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    key = torch.randint(0, 1 << 40, (num_rows,), dtype=torch.int64, device=device, generator=g)
    i32 = torch.randint(-(1 << 31), (1 << 31) - 1, (num_rows,), dtype=torch.int64, device=device,
                        generator=g).to(torch.int32)
    f64 = torch.randn(num_rows, dtype=torch.float64, device=device, generator=g)
    valid = torch.rand(num_rows, device=device, generator=g) >= 0.02
    vbits = _bitmap_device(valid)
    lens = torch.randint(8, 33, (num_rows,), dtype=torch.int32, device=device, generator=g)
    offs = torch.zeros(num_rows + 1, dtype=torch.int64, device=device)
    offs[1:] = torch.cumsum(lens, 0)
    total = int(offs[-1].item())
    assert total < (1 << 31), "utf8 data must fit int32 offsets"
    nb = num_rows // batch_rows
    # per-batch offsets starting at 0, as separate Arrow batches carry them
    idx = (torch.arange(nb, device=device)[:, None] * batch_rows +
           torch.arange(batch_rows + 1, device=device)[None, :])
    starts = offs[:-1:batch_rows]
    boffs = (offs[idx] - starts[:, None]).to(torch.int32).contiguous()
    bstart = starts.cpu().tolist()
    bend = offs[batch_rows::batch_rows].cpu().tolist()
    chars = torch.randint(97, 123, (total,), dtype=torch.uint8, device=device, generator=g)
    del offs, lens, idx
    kb, ib, fb, vb, cb = (t.view(torch.uint8) for t in (key, i32, f64, vbits, chars))
    batches = []
    for bi, b0 in enumerate(range(0, num_rows, batch_rows)):
        n = batch_rows
        bufs = [None, kb[b0 * 8:(b0 + n) * 8],
                None, ib[b0 * 4:(b0 + n) * 4],
                vb[b0 // 8:(b0 + n) // 8], fb[b0 * 8:(b0 + n) * 8],
                None, boffs[bi].view(torch.uint8), cb[bstart[bi]:bend[bi]]]
        batches.append(DeviceBatch(n, SPLIT_TYPES, bufs))
    nbytes = num_rows * (8 + 4 + 8) + nb * (batch_rows + 1) * 4 + num_rows // 8 + total
    return DeviceTable((key, i32, f64, vbits, boffs, chars), batches, num_rows, nbytes)


# ------------------------------------------------------------------ aggregate workload

def agg_host(num_rows: int, num_keys: int = 10000, batch_rows: int = BATCH_ROWS, seed: int = 42,
             null_frac: float = 0.01) -> List[HostBatch]:
    """int64 key over num_keys distinct values; float64 price ~ N(100, 25) with
    1% nulls; int32 qty ~ U[1, 50]."""
    keyset = _rng(10, seed).integers(-(1 << 62), 1 << 62, size=num_keys, dtype=np.int64)
    key = keyset[_rng(0, seed).integers(0, num_keys, size=num_rows)]
    r1 = _rng(1, seed)
    price = r1.normal(100.0, 25.0, size=num_rows)
    pvalid = r1.random(num_rows) >= null_frac
    qty = _rng(2, seed).integers(1, 51, size=num_rows).astype(np.int32)
    out = []
    for b0 in range(0, num_rows, batch_rows):
        b1 = min(num_rows, b0 + batch_rows)
        v = pvalid[b0:b1]
        out.append(HostBatch(b1 - b0, [
            HostColumn(T.INT64, key[b0:b1].copy()),
            HostColumn(T.DOUBLE, price[b0:b1].copy(), bitmap_from_bools(v) if not v.all() else None),
            HostColumn(T.INT32, qty[b0:b1].copy()),
        ]))
    return out


def device_from_host(batches: List[HostBatch], device: str = "cuda") -> List[DeviceBatch]:
    return [DeviceBatch.from_host(b, device) for b in batches]


def _agg_columns_device(g, n: int, key, device: str):
    """price f64 ~ N(100, 25) with 1% nulls, qty int32 ~ U[1, 50] next to a key."""
    import torch

    price = torch.randn(n, dtype=torch.float64, device=device, generator=g) * 25.0 + 100.0
    valid = torch.rand(n, device=device, generator=g) >= 0.01
    qty = torch.randint(1, 51, (n,), dtype=torch.int32, device=device, generator=g)
    return key, price, valid, qty


def agg_device(num_rows: int, num_keys: int = 10000, batch_rows: int = BATCH_ROWS, seed: int = 42,
               device: str = "cuda") -> DeviceTable:
    """The plain aggregate workload in HBM: int64 key over num_keys distinct
    values, price f64 ~ N(100, 25) with 1% nulls, qty int32 ~ U[1, 50]."""
    import torch

    assert num_rows % batch_rows == 0 and batch_rows % 64 == 0
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    keyset = torch.randint(-(1 << 62), 1 << 62, (num_keys,), dtype=torch.int64, device=device, generator=g)
    key = keyset[torch.randint(0, num_keys, (num_rows,), device=device, generator=g)]
    key, price, valid, qty = _agg_columns_device(g, num_rows, key, device)
    vbits = _bitmap_device(valid)
    kb, pb, vb, qb = (t.view(torch.uint8) for t in (key, price, vbits, qty))
    batches = []
    for b0 in range(0, num_rows, batch_rows):
        n = batch_rows
        bufs = [None, kb[b0 * 8:(b0 + n) * 8], vb[b0 // 8:(b0 + n) // 8], pb[b0 * 8:(b0 + n) * 8],
                None, qb[b0 * 4:(b0 + n) * 4]]
        batches.append(DeviceBatch(n, AGG_TYPES, bufs))
    nbytes = num_rows * (8 + 8 + 4) + num_rows // 8
    return DeviceTable((key, price, vbits, qty), batches, num_rows, nbytes)


def batched_agg_shapes(num_batches: int = 512, seed: int = 42, min_rows: int = 1000,
                       max_rows: int = 1000000, min_groups: int = 10, max_groups: int = 10 ** 7):
    """Row counts and group cardinalities, both log-uniform; groups capped at
    rows; row counts rounded up to multiples of 64 (whole validity words)."""
    r = _rng(20, seed)
    rows = np.exp(r.uniform(np.log(min_rows), np.log(max_rows), size=num_batches))
    rows = ((np.ceil(rows).astype(np.int64) + 63) // 64) * 64
    groups = np.exp(r.uniform(np.log(min_groups), np.log(max_groups), size=num_batches))
    groups = np.minimum(np.ceil(groups).astype(np.int64), rows)
    return rows, groups


def batched_agg_device(num_batches: int = 512, seed: int = 42, device: str = "cuda"):
    """The batched aggregate workload in HBM: one RecordBatch per independent
    GROUP BY (instance b), rows and cardinality log-uniform.  Returns the
    DeviceTable, the rows and the groups of every batch."""
    import torch

    rows, groups = batched_agg_shapes(num_batches, seed)
    total = int(rows.sum())
    g = torch.Generator(device=device)
    g.manual_seed(seed + 1)
    # key of row r of batch b: a mixed value of (b, u), u ~ U[0, groups_b)
    starts = np.zeros(num_batches + 1, np.int64)
    starts[1:] = np.cumsum(rows)
    gb = torch.from_numpy(np.repeat(groups, rows)).to(device)
    u = (torch.rand(total, dtype=torch.float64, device=device, generator=g) * gb).to(torch.int64)
    u = torch.minimum(u, gb - 1)
    bid = torch.from_numpy(np.repeat(np.arange(num_batches, dtype=np.int64), rows)).to(device)
    key = (u * 0x1E3779B97F4A7C15 + bid * 0x232BE59BD9B4E019) & 0x7FFFFFFFFFFFFFFF
    del u, bid, gb
    key, price, valid, qty = _agg_columns_device(g, total, key, device)
    vbits = _bitmap_device(valid)
    kb, pb, vb, qb = (t.view(torch.uint8) for t in (key, price, vbits, qty))
    batches = []
    for b in range(num_batches):
        b0, n = int(starts[b]), int(rows[b])
        bufs = [None, kb[b0 * 8:(b0 + n) * 8], vb[b0 // 8:(b0 + n) // 8], pb[b0 * 8:(b0 + n) * 8],
                None, qb[b0 * 4:(b0 + n) * 4]]
        batches.append(DeviceBatch(n, AGG_TYPES, bufs))
    nbytes = total * (8 + 8 + 4) + total // 8
    return DeviceTable((key, price, vbits, qty), batches, total, nbytes), rows, groups


# ------------------------------------------------------------------ join workload

def join_device(build_rows: int = 4 << 20, probe_rows: int = 256 << 20, zipf_s: float = 1.1,
                miss_frac: float = 0.05, batch_rows: int = BATCH_ROWS, seed: int = 42, device: str = "cuda"):
    """The hash-join workload in HBM.  Build: unique int64 keys in [0, 2^40)
    (an odd multiplier permutes that range), int32 and float64 payloads.  Probe:
    foreign keys build_key[perm[r]] with r ~ Zipf(s) over [1, build_rows]
    (inverse CDF), and miss_frac of them replaced by keys >= 2^40 that are absent.
    Returns (build DeviceTable, probe DeviceTable)."""
    import torch

    assert build_rows % batch_rows == 0 and probe_rows % batch_rows == 0
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    idx = torch.randperm(build_rows, device=device, generator=g).to(torch.int64)
    bkey = (idx * 0x9E3779B1 + 12345) & ((1 << 40) - 1)
    bi32 = torch.randint(-(1 << 31), (1 << 31) - 1, (build_rows,), dtype=torch.int64, device=device,
                         generator=g).to(torch.int32)
    bf64 = torch.randn(build_rows, dtype=torch.float64, device=device, generator=g)
    # Zipf ranks by inverse CDF
    w = torch.arange(1, build_rows + 1, dtype=torch.float64, device=device).pow(-zipf_s)
    cdf = torch.cumsum(w, 0)
    cdf /= cdf[-1].clone()
    del w
    perm = torch.randperm(build_rows, device=device, generator=g)
    fk = torch.empty(probe_rows, dtype=torch.int64, device=device)
    step = 32 << 20
    for a in range(0, probe_rows, step):
        b = min(probe_rows, a + step)
        u = torch.rand(b - a, dtype=torch.float64, device=device, generator=g)
        rank = torch.searchsorted(cdf, u).clamp_(max=build_rows - 1)
        k = bkey[perm[rank]]
        miss = torch.rand(b - a, device=device, generator=g) < miss_frac
        absent = (1 << 40) + torch.randint(0, 1 << 40, (b - a,), dtype=torch.int64, device=device, generator=g)
        fk[a:b] = torch.where(miss, absent, k)
        del u, rank, k, miss, absent
    del cdf, perm, idx
    build_types = [T.INT64, T.INT32, T.DOUBLE]
    kb, ib, fb = (t.view(torch.uint8) for t in (bkey, bi32, bf64))
    bb = []
    for b0 in range(0, build_rows, batch_rows):
        n = batch_rows
        bb.append(DeviceBatch(n, build_types, [None, kb[b0 * 8:(b0 + n) * 8], None, ib[b0 * 4:(b0 + n) * 4],
                                               None, fb[b0 * 8:(b0 + n) * 8]]))
    pk = fk.view(torch.uint8)
    pb = [DeviceBatch(batch_rows, [T.INT64], [None, pk[b0 * 8:(b0 + batch_rows) * 8]])
          for b0 in range(0, probe_rows, batch_rows)]
    build = DeviceTable((bkey, bi32, bf64), bb, build_rows, build_rows * 20)
    probe = DeviceTable((fk,), pb, probe_rows, probe_rows * 8)
    return build, probe
