"""The Splitter past its LDS-staged kernels' reach, bit-exact against the oracle.

More than 2048 partitions (the reference takes up to 65536, splitter.cc:218-221),
or per-tile counters that do not fit LDS (2048 partitions x 41 utf8 columns),
take the sorted path: every row's partition id, a stable radix sort of the rows
by id, and every column gathered in sorted order.  The same layout as the
LDS-staged path: partitions contiguous, rows of a partition in input order,
every partition's utf8 offsets starting at 0, bitmaps padded with ones."""
import numpy as np
import pytest

import oracle
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.batch import DeviceBatch, HostBatch, HostColumn
from gazelle_plugin_amd.splitter import SplitOptions, Splitter
from gazelle_plugin_amd._lib import Invalid
from splitcheck import assert_layout_equal, random_batches

pytestmark = pytest.mark.gpu

ALL_TYPES = [T.INT64, T.INT32, T.DOUBLE, T.STRING, T.BOOL, T.INT8, T.UINT16, T.DECIMAL128, T.NA, T.BINARY]


def run_gpu(ctx, batches, types, P, keys, mode="hash", batched=True, **opts):
    sp = Splitter.Make(ctx, mode, types, P, keys, SplitOptions(**opts))
    dev = [DeviceBatch.from_host(b) for b in batches]
    if batched:
        sp.SplitBatches(dev)
    else:
        for d in dev:
            sp.Split(d)
    sp.Stop()
    out = sp.read_partitions()
    sp.close()
    return out


def null_bits(bm, n):
    return int((np.unpackbits(np.asarray(bm).view(np.uint8), bitorder="little")[:n] == 0).sum())


@pytest.mark.parametrize("batched", [True, False])
@pytest.mark.parametrize("P", [2049, 4096, 65536])
def test_partition_counts_past_2048(ctx, P, batched):
    types = [T.INT64, T.STRING, T.DOUBLE]
    batches = random_batches(types, [4096, 700, 0, 5000], seed=P, null_frac=0.05)
    exp = oracle.split(batches, types, P, [0])
    assert_layout_equal(run_gpu(ctx, batches, types, P, [0], batched=batched), exp, types)


def test_all_types_ragged_with_null_counts(ctx):
    batches = random_batches(ALL_TYPES, [5000, 0, 1, 4096, 4097, 333], seed=31, null_frac=0.15)
    P = 5000
    exp = oracle.split(batches, ALL_TYPES, P, [0, 3])
    sp = Splitter.Make(ctx, "hash", ALL_TYPES, P, [0, 3])
    sp.SplitBatches([DeviceBatch.from_host(b) for b in batches])
    sp.Stop()
    assert_layout_equal(sp.read_partitions(), exp, ALL_TYPES)
    seg = sp.segment(0)
    N = int(exp["row_start"][-1])
    for c, t in enumerate(ALL_TYPES):
        v = exp["columns"][c].get("validity")
        if t != T.NA and v is not None:
            assert seg["columns"][c]["null_count"] == null_bits(v, N), c
    sp.close()


@pytest.mark.parametrize("batched", [True, False])
def test_round_robin_and_range(ctx, batched):
    types = [T.INT64, T.STRING, T.BOOL]
    batches = random_batches(types, [3000, 17, 4096, 5000], seed=41, null_frac=0.1)
    exp = oracle.split(batches, types, 3000, [], mode="rr")
    assert_layout_equal(run_gpu(ctx, batches, types, 3000, [], mode="rr", batched=batched), exp, types)
    rng = np.random.default_rng(42)
    ranged = [HostBatch(b.num_rows, [HostColumn(T.INT32, rng.integers(0, 4000, b.num_rows).astype(np.int32))]
                        + b.columns) for b in batches]
    rtypes = [T.INT32] + types
    exp = oracle.split(ranged, rtypes, 4000, [], mode="range")
    assert_layout_equal(run_gpu(ctx, ranged, rtypes, 4000, [], mode="range", batched=batched), exp, types)


def test_range_invalid_pid_raises(ctx):
    b = HostBatch(3, [HostColumn(T.INT32, np.array([0, 5000, 1], np.int32)), HostColumn(T.INT64, np.arange(3))])
    sp = Splitter.Make(ctx, "range", [T.INT32, T.INT64], 4000)
    with pytest.raises(Invalid):
        sp.Split(DeviceBatch.from_host(b))
    sp.close()


def test_row_ids_follow_rows(ctx):
    types = [T.INT64, T.DOUBLE]
    batches = random_batches(types, [3000, 4096, 17, 5000], seed=17, null_frac=0.1)
    P = 3000
    sp = Splitter.Make(ctx, "hash", types, P, [0], SplitOptions(append_row_ids=True))
    sp.SplitBatches([DeviceBatch.from_host(b) for b in batches])
    sp.Stop()
    lay = sp.read_partitions()
    exp = oracle.split(batches, types, P, [0])
    assert_layout_equal(lay, exp, types)
    keys_in = np.concatenate([b.columns[0].values for b in batches])
    rid = lay["columns"][2]["values"].view(np.int32)
    assert len(rid) == sum(b.num_rows for b in batches)
    for p in np.flatnonzero(np.diff(lay["row_start"])):
        a, b = lay["row_start"][p], lay["row_start"][p + 1]
        assert np.all(np.diff(rid[a:b]) > 0)
        assert np.array_equal(keys_in[rid[a:b]], lay["columns"][0]["values"].view(np.int64)[a:b])
    sp.close()


def test_segment_batch_counts_follow_the_input_batches(ctx):
    types = [T.INT64, T.STRING]
    batches = random_batches(types, [5000, 100, 4096, 3000], seed=23, null_frac=0.05)
    P = 2500
    sp = Splitter.Make(ctx, "hash", types, P, [0])
    sp.SplitBatches([DeviceBatch.from_host(b) for b in batches])
    sp.Stop()
    got = np.asarray(sp.segment_batch_counts(0)).reshape(len(batches), P)
    for bi, b in enumerate(batches):
        pids = np.asarray(oracle.split_pids([b], types, P, [0]))
        assert np.array_equal(got[bi], np.bincount(pids, minlength=P)), bi
    sp.close()


def test_counters_past_the_lds_budget(ctx):
    """2048 partitions x (1 + 40 utf8 columns) counters do not fit LDS: sorted path."""
    types = [T.INT64] + [T.STRING] * 40
    batches = random_batches(types, [1500, 900], seed=43, null_frac=0.1)
    exp = oracle.split(batches, types, 2048, [0])
    assert_layout_equal(run_gpu(ctx, batches, types, 2048, [0]), exp, types)


def test_batches_sharing_one_utf8_buffer(ctx):
    """utf8 slices of one device buffer: the data output sized from the scanned total."""
    import torch

    types = [T.INT64, T.STRING]
    batches = random_batches(types, [3000, 4096, 17, 5000, 1], seed=5, null_frac=0.1)
    P = 3333
    exp = oracle.split(batches, types, P, [0])
    shared = torch.from_numpy(np.concatenate([b.columns[1].values for b in batches]).astype(np.uint8)).to("cuda")
    dev, base = [], 0
    for b in batches:
        col = b.columns[1]
        offs = (col.offsets.astype(np.int64) + base).astype(np.int32)
        base += len(col.values)
        tensors = list(DeviceBatch.from_host(b).tensors)
        tensors[3] = torch.from_numpy(offs.view(np.uint8).copy()).to("cuda")
        tensors[4] = shared
        dev.append(DeviceBatch(b.num_rows, types, tensors))
    sp = Splitter.Make(ctx, "hash", types, P, [0])
    sp.SplitBatches(dev)
    sp.Stop()
    assert_layout_equal(sp.read_partitions(), exp, types)
    sp.close()


@pytest.mark.parametrize("P", [200, 5000])
def test_utf8_offsets_past_the_declared_data_size(ctx, P):
    """A batch declaring fewer data bytes than its offsets span (the device
    buffer itself holds them all): Stop reports Invalid, and no string is
    written past the data output sized from the declared sizes."""
    types = [T.INT64, T.STRING]
    batches = random_batches(types, [4096, 3000], seed=7, null_frac=0.0)
    dev = [DeviceBatch.from_host(b) for b in batches]
    dev[1].sizes[4] = 8  # the utf8 data buffer: 8 bytes declared
    sp = Splitter.Make(ctx, "hash", types, P, [0])
    sp.SplitBatches(dev)
    with pytest.raises(Invalid, match="declared data"):
        sp.Stop()
    sp.close()
