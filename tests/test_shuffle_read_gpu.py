"""The shuffle read on the GPU: a Splitter's data file, read back per partition
into device batches that point into the file's device copy, feeds the next
device stage unchanged.  A second (single-partition) split of each partition's
batches reproduces the oracle's rows of that partition bit for bit, and a hash
aggregate over the shuffled batches equals the oracle's aggregate of the same
rows."""
import numpy as np
import pyarrow as pa
import pytest

import oracle
from gazelle_plugin_amd import datagen
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.aggregate import AggOptions, HashAggregateKernel, parse_action
from gazelle_plugin_amd.batch import DeviceBatch, HostBatch, HostColumn, bitmap_from_bools, bools_from_bitmap
from gazelle_plugin_amd.shuffle_reader import ShuffleReader
from gazelle_plugin_amd.splitter import SplitOptions, Splitter
from aggcheck import CONFIG0_ACTIONS, assert_results_equal
from splitcheck import assert_layout_equal, random_batches

pytestmark = pytest.mark.gpu

ARROW = {T.INT64: pa.int64(), T.INT32: pa.int32(), T.DOUBLE: pa.float64(), T.STRING: pa.string(),
         T.BOOL: pa.bool_(), T.INT8: pa.int8(), T.UINT16: pa.uint16(), T.NA: pa.null(), T.BINARY: pa.binary(),
         T.DATE32: pa.date32()}


def _bits(bm, start, n):
    return np.unpackbits(np.asarray(bm, np.uint8).reshape(-1).view(np.uint8),
                         bitorder="little")[start:start + n].astype(bool)


def partition_rows(exp, types, p) -> HostBatch:
    """Partition p of a concatenated partition layout, as one host batch."""
    r0, r1 = int(exp["row_start"][p]), int(exp["row_start"][p + 1])
    n = r1 - r0
    cols = []
    for c, t in enumerate(types):
        col = exp["columns"][c]
        if t == T.NA:
            cols.append(HostColumn(T.NA, None))
            continue
        valid = None if col["validity"] is None else bitmap_from_bools(_bits(col["validity"], r0, n))
        if T.is_binary(t):
            off = np.asarray(col["offsets"], np.int64)[r0 + p: r1 + p + 1]
            bs = int(col["byte_start"][p])
            data = np.asarray(col["data"], np.uint8)[bs + off[0]: bs + off[-1]].copy()
            cols.append(HostColumn(t, data, valid, (off - off[0]).astype(np.int32)))
        elif t == T.BOOL:
            cols.append(HostColumn(t, bitmap_from_bools(_bits(col["values"], r0, n)), valid))
        else:
            w = T.width(t)
            raw = np.asarray(col["values"]).view(np.uint8).reshape(-1)[r0 * w: r1 * w].copy()
            cols.append(HostColumn(t, raw.view(T.numpy_dtype(t)), valid))
    return HostBatch(n, cols)


def shuffle(ctx, batches, types, P, keys, buffer_size):
    """Hash split + data file; returns (file bytes, partition lengths)."""
    sp = Splitter.Make(ctx, "hash", types, P, keys, SplitOptions(buffer_size=buffer_size))
    sp.SplitBatches([DeviceBatch.from_host(b) for b in batches])
    sp.Stop()
    schema = pa.schema([(f"c{i}", ARROW[t]) for i, t in enumerate(types)])
    data, lengths = sp.ipc_stream(schema.serialize().to_pybytes())
    sp.close()
    return data, lengths


@pytest.mark.parametrize("batched", [True, False])
def test_read_back_then_split_again(ctx, batched):
    types = [T.INT64, T.INT32, T.DOUBLE, T.STRING, T.BOOL, T.INT8, T.UINT16, T.NA, T.BINARY, T.DATE32]
    batches = random_batches(types, [5000, 0, 1, 4096, 777, 3000], seed=31, null_frac=0.15)
    P = 7
    exp = oracle.split(batches, types, P, [0, 3])
    data, lengths = shuffle(ctx, batches, types, P, [0, 3], buffer_size=500)
    parts = ShuffleReader(types).read_data_file(data, lengths)
    assert len(parts) == P
    for p in range(P):
        n = int(exp["row_start"][p + 1] - exp["row_start"][p])
        assert sum(b.num_rows for b in parts[p]) == n
        if n == 0:
            assert parts[p] == []
            continue
        sp = Splitter.Make(ctx, "single", types, 1)
        if batched:
            sp.SplitBatches(parts[p])
        else:
            for b in parts[p]:
                sp.Split(b)
        sp.Stop()
        got = sp.read_partitions()
        sp.close()
        want = oracle.split([partition_rows(exp, types, p)], types, 1, [], mode="single")
        assert_layout_equal(got, want, types)


def test_aggregate_over_shuffled_batches(ctx):
    batches = datagen.agg_host(60000, num_keys=5000, batch_rows=7000)
    types = datagen.AGG_TYPES
    P = 5
    exp = oracle.split(batches, types, P, [0])
    data, lengths = shuffle(ctx, batches, types, P, [0], buffer_size=4096)
    parts = ShuffleReader(types).read_data_file(data, lengths)
    actions = [parse_action(n, i) for n, i in CONFIG0_ACTIONS]
    for p in range(P):
        k = HashAggregateKernel.Make(ctx, types, CONFIG0_ACTIONS, AggOptions())
        it = k.MakeResultIterator()
        it.ProcessBatches(parts[p])
        got = it.results()
        it.close()
        assert_results_equal(got, oracle.aggregate([partition_rows(exp, types, p)], types, actions), 1e-12)


def test_one_partition_stream_uploaded_alone(ctx):
    types = [T.INT64, T.STRING]
    batches = random_batches(types, [3000, 2000], seed=32, null_frac=0.1)
    P = 3
    exp = oracle.split(batches, types, P, [0])
    data, lengths = shuffle(ctx, batches, types, P, [0], buffer_size=256)
    off = np.concatenate([[0], np.cumsum(lengths)])
    for p in range(P):
        part = data[off[p]:off[p + 1]]
        if len(part) == 0:
            continue
        got_batches = ShuffleReader(types).read(part)
        sp = Splitter.Make(ctx, "single", types, 1)
        sp.SplitBatches(got_batches)
        sp.Stop()
        got = sp.read_partitions()
        sp.close()
        assert_layout_equal(got, oracle.split([partition_rows(exp, types, p)], types, 1, [], mode="single"), types)
