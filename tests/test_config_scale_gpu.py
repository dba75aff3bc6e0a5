"""The workloads at their full sizes, checked completely (not sampled):

- the shuffle split of 64M rows into 200 partitions: with the row-id column,
  every output row is its input row (all columns bit for bit), the rows of a
  partition are exactly the input rows whose oracle partition id is that
  partition, in input order;
- the hash join of a 4M-row build with a 256M-row Zipf probe: the semi join's
  output is the probe keys present in the build side, in probe order, the anti
  join's the absent ones (together: every probe row), and the inner join's rows
  are the semi join's with each key's build row (build columns bit for bit);
- the aggregate query over 8M rows and 10k groups, and the 512-instance
  aggregate set (72M rows, 21M groups), against the oracle's aggregate.

Expected values come from the inputs themselves (torch on the device) and the
oracle's partition ids and aggregates."""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from gazelle_plugin_amd import _lib as L
from gazelle_plugin_amd import datagen
from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.aggregate import HashAggregateKernel, HashAggregateResultIterator, parse_action
from gazelle_plugin_amd.batch import HostBatch, HostColumn
from gazelle_plugin_amd.hip import D2D, hip
from gazelle_plugin_amd.join import ConditionedProbeKernel, HashRelationKernel
from gazelle_plugin_amd.splitter import SplitOptions, Splitter, make_batch_array
from aggcheck import CONFIG0_ACTIONS, assert_results_equal

pytestmark = pytest.mark.gpu


def device_tensor(ctx, ptr: int, nbytes: int, dtype) -> torch.Tensor:
    """A torch copy of nbytes of library-owned device memory (the context's
    stream drained first)."""
    ctx.synchronize()
    out = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device="cuda")
    if nbytes > 0:
        torch.cuda.synchronize()
        rc = hip().hipMemcpy(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ptr), int(nbytes), D2D)
        assert rc == 0, rc
    return out[:nbytes].view(dtype)


def bits_of(bitmap: torch.Tensor, n: int) -> torch.Tensor:
    """Bools of the first n bits of an LSB-first bitmap."""
    shifts = torch.arange(8, device=bitmap.device, dtype=torch.uint8)
    return ((bitmap[: (n + 7) // 8, None] >> shifts) & 1).reshape(-1)[:n].bool()


def test_split_64m_rows_complete(ctx):
    P = 200
    tab = datagen.split_device(64 << 20)
    N = tab.num_rows
    key, i32, f64, vbits, boffs, chars = tab.tensors
    sp = Splitter.Make(ctx, "hash", datagen.SPLIT_TYPES, P, [0], SplitOptions(append_row_ids=True))
    sp.SplitBatches(tab.batches, make_batch_array(tab.batches))
    sp.Stop()
    assert sp.num_segments() == 1
    rs = np.zeros(P + 1, np.int64)
    L.check(L.lib().gz_split_segment_rows(sp._h, 0, rs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
    assert rs[-1] == N

    def column(c):
        vals, valid, offs, data = (ctypes.c_void_p() for _ in range(4))
        bs = np.zeros(P + 1, np.int64)
        nulls = ctypes.c_int64()
        L.check(L.lib().gz_split_segment_column(sp._h, 0, c, ctypes.byref(vals), ctypes.byref(valid),
                                                ctypes.byref(offs), ctypes.byref(data),
                                                bs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                ctypes.byref(nulls)))
        return vals.value, valid.value, offs.value, data.value, bs, nulls.value

    # the oracle's partition id of every input row
    kh = key.cpu().numpy()
    pid = torch.from_numpy(oracle.split_pids([HostBatch(N, [HostColumn(T.INT64, kh)])], [T.INT64], P, [0])).cuda()
    del kh
    counts = torch.bincount(pid.long(), minlength=P).cpu().numpy()
    np.testing.assert_array_equal(np.diff(rs), counts)
    # row ids: inside a partition, increasing input rows of that partition
    rid = device_tensor(ctx, column(4)[0], 4 * N, torch.int32).long()
    part = torch.repeat_interleave(torch.arange(P, device="cuda"), torch.from_numpy(np.diff(rs)).cuda())
    assert torch.equal(pid[rid].long(), part)
    step = rid[1:] - rid[:-1]
    inside = part[1:] == part[:-1]
    assert bool((step[inside] > 0).all())
    del step, inside, part, pid
    # every column of every output row is its input row's
    assert torch.equal(device_tensor(ctx, column(0)[0], 8 * N, torch.int64), key[rid])
    assert torch.equal(device_tensor(ctx, column(1)[0], 4 * N, torch.int32), i32[rid])
    v, valid, _, _, _, nulls = column(2)
    assert torch.equal(device_tensor(ctx, v, 8 * N, torch.int64), f64.view(torch.int64)[rid])
    in_ok = bits_of(vbits, N)
    assert torch.equal(bits_of(device_tensor(ctx, valid, (N + 7) // 8, torch.uint8), N), in_ok[rid])
    assert nulls == int((~in_ok).sum())
    del in_ok
    _, _, offs, data, bs, _ = column(3)
    ooff = device_tensor(ctx, offs, 4 * (N + P), torch.int32).long()
    odata = device_tensor(ctx, data, int(bs[-1]), torch.uint8)
    # input lengths and global starts of every row
    ilen = (boffs[:, 1:] - boffs[:, :-1]).reshape(-1).long()
    istart = torch.zeros(N + 1, dtype=torch.long, device="cuda")
    istart[1:] = torch.cumsum(ilen, 0)
    # output lengths: partition p's offsets sit at [rs[p] + p, rs[p+1] + p]
    starts = torch.from_numpy(rs[:-1] + np.arange(P)).cuda()
    pos = torch.arange(N, device="cuda") + torch.repeat_interleave(torch.arange(P, device="cuda"),
                                                                   torch.from_numpy(np.diff(rs)).cuda())
    olen = ooff[pos + 1] - ooff[pos]
    assert torch.equal(olen, ilen[rid])
    assert bool((ooff[starts] == 0).all())
    # bytes: output row i's string at byte_start[p] + offset equals its input string
    obeg = torch.from_numpy(bs[:-1]).cuda()[torch.repeat_interleave(torch.arange(P, device="cuda"),
                                                                     torch.from_numpy(np.diff(rs)).cuda())] + ooff[pos]
    assert bool((obeg[1:] == obeg[:-1] + olen[:-1]).all())  # the output strings are packed in row order
    src = torch.repeat_interleave(istart[:-1][rid] - obeg, olen)  # input byte index minus output byte index
    assert torch.equal(odata, chars[torch.arange(odata.numel(), device="cuda") + src])
    sp.close()


def _join_columns(ctx, probe, n, ncols):
    cols = []
    for c in range(ncols):
        vals, valid, offs, data = (ctypes.c_void_p() for _ in range(4))
        L.check(L.lib().gz_probe_result_column(probe._h, c, ctypes.byref(vals), ctypes.byref(valid),
                                               ctypes.byref(offs), ctypes.byref(data)))
        cols.append((vals.value, valid.value))
    return cols


def test_zipf_join_4m_build_256m_probe_complete(ctx):
    build, probe = datagen.join_device()
    bkey, bi32, bf64 = build.tensors
    (fk,) = probe.tensors
    rel = HashRelationKernel.Make(ctx, [T.INT64, T.INT32, T.DOUBLE], 0)
    rel.EvaluateBatches(build.batches)
    rel.Finish()
    assert rel.num_keys() == build.num_rows
    parr = make_batch_array(probe.batches)
    present = torch.isin(fk, bkey)
    exp_semi = fk[present]
    exp_anti = fk[~present]
    del present
    out = {}
    for jt in ("semi", "anti", "inner"):
        result = [(1, 0), (0, 1), (0, 2)] if jt == "inner" else [(1, 0)]
        pk = ConditionedProbeKernel.Make(rel, [T.INT64], 0, jt, False, result)
        n, per = pk.ProcessBatches(probe.batches, parr)
        assert int(per.sum()) == n
        cols = _join_columns(ctx, pk, n, len(result))
        keys = device_tensor(ctx, cols[0][0], 8 * n, torch.int64)
        if jt == "inner":
            # the build row of every output key: its position among the sorted build keys
            sk, order = torch.sort(bkey)
            row = order[torch.searchsorted(sk, keys)]
            assert torch.equal(bkey[row], keys)
            assert torch.equal(device_tensor(ctx, cols[1][0], 4 * n, torch.int32), bi32[row])
            assert torch.equal(device_tensor(ctx, cols[2][0], 8 * n, torch.int64), bf64.view(torch.int64)[row])
            for c in (1, 2):  # build columns without nulls: every output row valid
                assert bool((device_tensor(ctx, cols[c][1], n, torch.uint8) == 1).all())
            del sk, order, row
        out[jt] = keys
        pk.close()
    assert torch.equal(out["semi"], exp_semi)
    assert torch.equal(out["anti"], exp_anti)
    assert out["semi"].numel() + out["anti"].numel() == probe.num_rows
    assert torch.equal(out["inner"], exp_semi)  # unique build keys: one row per matching probe row
    rel.close()


def _host_batches(tab, types, rows):
    """Host copies of a device table's batches (one per entry of rows)."""
    key, price, vbits, qty = (t.cpu().numpy() for t in tab.tensors)
    vall = np.unpackbits(vbits, bitorder="little")
    out, r0 = [], 0
    for n in rows:
        n = int(n)
        vb = vall[r0:r0 + n]
        out.append(HostBatch(n, [HostColumn(T.INT64, key[r0:r0 + n]),
                                 HostColumn(T.DOUBLE, price[r0:r0 + n], np.packbits(vb, bitorder="little")),
                                 HostColumn(T.INT32, qty[r0:r0 + n])]))
        r0 += n
    return out


def test_aggregate_query_8m_rows_matches_oracle(ctx):
    tab = datagen.agg_device(8 << 20)
    k = HashAggregateKernel.Make(ctx, datagen.AGG_TYPES, CONFIG0_ACTIONS)
    it = k.MakeResultIterator()
    it.ProcessBatches(tab.batches)
    got = it.results()
    it.close()
    host = _host_batches(tab, datagen.AGG_TYPES, [tab.num_rows])
    exp = oracle.aggregate(host, datagen.AGG_TYPES, [parse_action(n, i) for n, i in CONFIG0_ACTIONS])
    assert len(got[0][1]) == 10000
    assert_results_equal(got, exp, 1e-12)


def test_aggregate_set_512_instances_matches_oracle(ctx):
    tab, rows, groups = datagen.batched_agg_device(512)
    k = HashAggregateKernel.Make(ctx, datagen.AGG_TYPES, CONFIG0_ACTIONS)
    it = HashAggregateResultIterator(k, num_instances=len(tab.batches))
    arr = make_batch_array(tab.batches)
    it.Autotune(tab.batches, list(range(len(tab.batches))), arr)
    it.Capture()
    it.Launch()
    ctx.synchronize()
    acts = [parse_action(n, i) for n, i in CONFIG0_ACTIONS]
    host = _host_batches(tab, datagen.AGG_TYPES, rows)
    for i, hb in enumerate(host):
        exp = oracle.aggregate([hb], datagen.AGG_TYPES, acts)
        got = it.results(i)
        assert len(got[0][1]) == len(exp[0][1]), (i, rows[i], groups[i])
        assert_results_equal(got, exp, 1e-12)
    it.close()
