"""Parity of the HIP Splitter with the oracle (bit-exact), through the C ABI."""
import numpy as np
import pytest

import oracle
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd import datagen
from gazelle_plugin_amd.batch import DeviceBatch, HostBatch, HostColumn
from gazelle_plugin_amd.splitter import SplitOptions, Splitter
from gazelle_plugin_amd._lib import GazelleError, Invalid
from golden import fixtures as F
from splitcheck import ordinary_float_keys, assert_layout_equal, random_batches

pytestmark = pytest.mark.gpu


def run_gpu(ctx, batches, types, P, keys, mode="hash", batched=True, tile_rows=0):
    sp = Splitter.Make(ctx, mode, types, P, keys, SplitOptions(tile_rows=tile_rows))
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


@pytest.mark.parametrize("batched", [True, False])
def test_split_workload_shape_bit_exact(ctx, batched):
    batches = datagen.split_host(3 * 4096 + 1000)
    types = datagen.SPLIT_TYPES
    exp = oracle.split(batches, types, 200, [0])
    got = run_gpu(ctx, batches, types, 200, [0], batched=batched)
    assert_layout_equal(got, exp, types)


@pytest.mark.parametrize("tile_rows", [1024, 2048, 4096])
def test_split_all_types_ragged(ctx, tile_rows):
    types = [T.INT64, T.INT32, T.DOUBLE, T.STRING, T.BOOL, T.INT8, T.UINT16, T.DECIMAL128, T.NA, T.BINARY]
    batches = random_batches(types, [5000, 0, 1, 4096, 4097, 333], seed=11, null_frac=0.15)
    exp = oracle.split(batches, types, 37, [0, 3])
    got = run_gpu(ctx, batches, types, 37, [0, 3], tile_rows=tile_rows)
    assert_layout_equal(got, exp, types)


@pytest.mark.parametrize("P", [1, 2, 255, 256, 257, 2048])
def test_split_partition_counts(ctx, P):
    types = [T.INT64, T.STRING, T.DOUBLE]
    batches = random_batches(types, [4096, 700], seed=P, null_frac=0.05)
    exp = oracle.split(batches, types, P, [0])
    got = run_gpu(ctx, batches, types, P, [0])
    assert_layout_equal(got, exp, types)


@pytest.mark.parametrize("key_type", [T.INT32, T.INT16, T.INT8, T.BOOL, T.DATE32, T.DOUBLE, T.FLOAT, T.STRING])
def test_split_key_types_match_oracle(ctx, key_type):
    types = [key_type, T.INT64]
    batches = ordinary_float_keys(random_batches(types, [3000, 123], seed=key_type, null_frac=0.1), [0])
    exp = oracle.split(batches, types, 200, [0])
    got = run_gpu(ctx, batches, types, 200, [0])
    assert_layout_equal(got, exp, types)


@pytest.mark.parametrize("case", [0, 1, 2])
def test_reference_fixtures_on_gpu(ctx, case):
    fx = F.load("split_fixtures.json")
    test = fx["tests"][case]
    schema = fx["schema"]
    named = {k: F.batch(schema, v) for k, v in fx["batches"].items()}
    inputs = [named[k] for k in test["split"]]
    types = [F.type_id(t) for _, t in schema]
    mode = test["short_name"]
    if mode == "range":
        inputs = [HostBatch(b.num_rows, [HostColumn(T.INT32, np.array(test["pids"][k], np.int32))] + b.columns)
                  for b, k in zip(inputs, test["split"])]
        types = [T.INT32] + types
    P = test["num_partitions"]
    exp = oracle.split(inputs, types, P, [], mode=mode)
    out_types = types[1:] if mode == "range" else types
    for batched in (True, False):
        got = run_gpu(ctx, inputs, types, P, [], mode=mode, batched=batched)
        assert_layout_equal(got, exp, out_types)


def test_range_invalid_pid_raises(ctx):
    b = HostBatch(3, [HostColumn(T.INT32, np.array([0, 5, 1], np.int32)), HostColumn(T.INT64, np.arange(3))])
    sp = Splitter.Make(ctx, "range", [T.INT32, T.INT64], 2)
    with pytest.raises(Invalid):
        sp.Split(DeviceBatch.from_host(b))
    sp.close()


def test_unknown_partitioning_not_implemented(ctx):
    with pytest.raises(GazelleError) as ei:
        Splitter.Make(ctx, "roundrobin2", [T.INT64], 2)
    assert ei.value.code == 10


def test_hbm_budget_out_of_memory():
    from gazelle_plugin_amd.context import Context
    from gazelle_plugin_amd._lib import OutOfMemory

    small = Context(0, hbm_budget=1 << 16)
    batches = datagen.split_host(20000)
    sp = Splitter.Make(small, "hash", datagen.SPLIT_TYPES, 200, [0])
    dev = [DeviceBatch.from_host(b) for b in batches]
    with pytest.raises(OutOfMemory):
        sp.SplitBatches(dev)
    sp.close()
    small.close()


def test_device_generated_full_path_properties(ctx):
    """Size-independent properties on a device-generated input (no oracle run):
    partitions concatenate to a permutation of the input, counts match."""
    import torch

    tab = datagen.split_device(64 * 4096)
    sp = Splitter.Make(ctx, "hash", datagen.SPLIT_TYPES, 200, [0])
    sp.SplitBatches(tab.batches)
    sp.Stop()
    seg = sp.segment(0)
    key = tab.tensors[0].cpu().numpy()
    gk = seg["columns"][0]["values"].view(np.int64)
    assert np.array_equal(np.sort(gk), np.sort(key))
    # partition of every output key equals its run
    pids = np.array([oracle.pmod(oracle.hash_long(int(k)), 200) for k in gk[:2000]])
    rs = seg["row_start"]
    run = np.searchsorted(rs, np.arange(2000), side="right") - 1
    assert np.array_equal(pids, run)
    nulls = int((np.unpackbits(tab.tensors[3].cpu().numpy(), bitorder="little") == 0).sum())
    assert seg["columns"][2]["null_count"] == nulls
    sp.close()
    torch.cuda.synchronize()


def test_row_ids_follow_rows(ctx):
    """The optional row-index column lands next to its row: within every
    partition the indices ascend (stable split) and index the input rows."""
    from gazelle_plugin_amd.splitter import SplitOptions, Splitter

    types = [T.INT64, T.DOUBLE]
    batches = random_batches(types, [3000, 4096, 17, 5000], seed=17, null_frac=0.1)
    sp = Splitter.Make(ctx, "hash", types, 37, [0], SplitOptions(append_row_ids=True))
    sp.SplitBatches([DeviceBatch.from_host(b) for b in batches])
    sp.Stop()
    lay = sp.read_partitions()
    exp = oracle.split(batches, types, 37, [0], "hash", 42)
    keys_in = np.concatenate([b.columns[0].values for b in batches])
    rid = lay["columns"][2]["values"].view(np.int32)
    assert len(rid) == sum(b.num_rows for b in batches)
    for p in range(37):
        a, b = lay["row_start"][p], lay["row_start"][p + 1]
        assert np.all(np.diff(rid[a:b]) > 0)
        got_keys = lay["columns"][0]["values"].view(np.int64)[a:b]
        assert np.array_equal(keys_in[rid[a:b]], got_keys)
    assert np.array_equal(lay["row_start"], exp["row_start"])
    sp.close()


def test_segment_batch_counts_follow_the_input_batches(ctx):
    """Rows per (batch, partition) of a batched call, from the tile counters."""
    types = [T.INT64, T.STRING]
    batches = random_batches(types, [5000, 100, 4096, 3000], seed=23, null_frac=0.05)
    sp = Splitter.Make(ctx, "hash", types, 31, [0])
    sp.SplitBatches([DeviceBatch.from_host(b) for b in batches])
    sp.Stop()
    got = np.asarray(sp.segment_batch_counts(0)).reshape(len(batches), 31)
    for bi, b in enumerate(batches):
        pids = np.asarray(oracle.split_pids([b], types, 31, [0]))
        assert np.array_equal(got[bi], np.bincount(pids, minlength=31)), bi
    sp.close()


def test_batches_sharing_one_utf8_buffer(ctx):
    """Batches whose utf8 bytes are slices of one device buffer (offsets that do
    not start at 0, the same data address in every batch): the output bytes are
    sized from the scanned totals, and the layout equals the oracle's."""
    import torch

    types = [T.INT64, T.STRING]
    batches = random_batches(types, [3000, 4096, 17, 5000, 1], seed=5, null_frac=0.1)
    exp = oracle.split(batches, types, 64, [0])
    shared = torch.from_numpy(np.concatenate([b.columns[1].values for b in batches]).astype(np.uint8)).to("cuda")
    dev, base = [], 0
    for b in batches:
        col = b.columns[1]
        offs = (col.offsets.astype(np.int64) + base).astype(np.int32)
        base += len(col.values)
        host = DeviceBatch.from_host(b)
        tensors = list(host.tensors)
        tensors[3] = torch.from_numpy(offs.view(np.uint8).copy()).to("cuda")
        tensors[4] = shared
        dev.append(DeviceBatch(b.num_rows, types, tensors))
    sp = Splitter.Make(ctx, "hash", types, 64, [0])
    sp.SplitBatches(dev)
    sp.Stop()
    got = sp.read_partitions()
    sp.close()
    assert_layout_equal(got, exp, types)


@pytest.mark.parametrize("key_type", [T.DOUBLE, T.FLOAT])
@pytest.mark.parametrize("bits", ["neg_zero", "nan"])
def test_split_refuses_unpinned_float_keys(ctx, key_type, bits):
    """-0.0 and NaN hash differently across Spark versions: Stop reports
    NotImplemented, as the oracle refuses them."""
    from gazelle_plugin_amd import _lib as L

    types = [key_type, T.INT64]
    batches = ordinary_float_keys(random_batches(types, [500], seed=3, null_frac=0.0), [0])
    vals = batches[0].columns[0].values
    u = vals.view(np.uint64 if key_type == T.DOUBLE else np.uint32)
    u[123] = (1 << 63 if key_type == T.DOUBLE else 1 << 31) if bits == "neg_zero" else \
        (0x7FF8000000000001 if key_type == T.DOUBLE else 0x7FC00001)
    with pytest.raises(ValueError):
        oracle.split(batches, types, 200, [0])
    sp = Splitter.Make(ctx, "hash", types, 200, [0])
    sp.SplitBatches([DeviceBatch.from_host(b) for b in batches])
    with pytest.raises(L.NotImplementedStatus):
        sp.Stop()
    sp.close()


@pytest.mark.parametrize("batched", [True, False])
def test_split_wide_schema_many_keys(ctx, batched):
    """A wide row (128 columns: every fixed width, bool, utf8 and binary, nulls)
    hashed on 16 key columns of every hashable type, split into 200
    partitions, bit-exact with the oracle; per-batch and batched."""
    base = [T.INT64, T.INT32, T.DOUBLE, T.STRING, T.BOOL, T.INT8, T.UINT16, T.DECIMAL128, T.INT16, T.FLOAT,
            T.DATE32, T.BINARY, T.UINT64, T.UINT8, T.UINT32, T.NA]
    types = (base * 8)[:128]
    keys = [0, 1, 3, 4, 5, 8, 10, 16, 17, 19, 20, 21, 24, 26, 35, 40]
    assert len(keys) == 16 and all(types[k] in (T.INT64, T.INT32, T.STRING, T.BOOL, T.INT8, T.INT16, T.DATE32)
                                   for k in keys)
    batches = random_batches(types, [3000, 1, 4096, 257], seed=29, null_frac=0.1)
    exp = oracle.split(batches, types, 200, keys)
    got = run_gpu(ctx, batches, types, 200, keys, batched=batched)
    assert_layout_equal(got, exp, types)


def test_split_limits(ctx):
    """Beyond the device splitter's limits: NotImplemented, not a wrong split."""
    from gazelle_plugin_amd import _lib as L

    with pytest.raises(L.NotImplementedStatus):
        Splitter.Make(ctx, "hash", [T.INT64] * 129, 200, [0])
    with pytest.raises(L.NotImplementedStatus):
        Splitter.Make(ctx, "hash", [T.INT64] * 20, 200, list(range(17)))
    with pytest.raises(L.Invalid):  # the reference's limit (splitter.cc:218-221)
        Splitter.Make(ctx, "hash", [T.INT64, T.STRING], 65537, [0])
