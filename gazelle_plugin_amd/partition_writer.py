"""Host side of the shuffle writer: the reference's output record-batch cuts.

The device split keeps every partition as one contiguous run per segment.  The
reference emits each partition as a sequence of record batches whose boundaries
come from its partition-buffer policy in ``Splitter::DoSplit``
(cpp/src/shuffle/splitter.cc:889-925): a partition buffer holds
``max(CalculateSplitBatchSize, rows of this input batch)`` rows; when the rows of
the next input batch do not fit, the filled buffer is cached as one record batch
(``CacheRecordBatch``, :440-608) and the buffer is reused or reallocated.  These
functions reproduce those cuts from the per-input-batch row counts, so the host
can slice the device runs into the same record batches.
"""
from __future__ import annotations

from typing import List, Sequence


def split_batch_size(buffer_size: int, offheap_per_task: int = 0, size_per_row: int = 0,
                     num_partitions: int = 1) -> int:
    """``Splitter::CalculateSplitBatchSize`` (splitter.cc:800-853)."""
    if offheap_per_task > 0 and size_per_row > 0:
        pre = (offheap_per_task // size_per_row // num_partitions) >> 2
    else:
        pre = buffer_size
    return min(pre, buffer_size)


def reference_batch_rows(counts: Sequence[int], buffer_size: int, prefer_spill: bool = True,
                         batch_size: int = 0) -> List[int]:
    """Row counts of the record batches the reference writes for one partition,
    given that partition's row count in each input batch, in order."""
    bs = batch_size or buffer_size
    size = 0  # partition_buffer_size_[pid]
    base = 0  # partition_buffer_idx_base_[pid]
    out: List[int] = []
    for cnt in counts:
        cnt = int(cnt)
        if cnt <= 0:
            continue
        new_size = max(bs, cnt)
        if size == 0:
            size = new_size
        elif base + cnt > size:
            if base > 0:
                out.append(base)  # CacheRecordBatch
            base = 0
            if prefer_spill:
                if new_size > size:
                    size = new_size
            else:
                size = new_size
        base += cnt
    if base > 0:
        out.append(base)  # Stop(): CacheRecordBatch of every partition
    return out


def record_batch_message(num_rows: int, nodes, buffers, body_length: int) -> bytes:
    """Encapsulated Arrow IPC message of a record batch (the metadata the device
    partition writer puts before each body), from the C ABI's host encoder:
    nodes = [(length, null_count)], buffers = [(offset, length)]."""
    import ctypes

    import numpy as np

    from . import _lib as L

    nd = np.asarray(nodes, dtype=np.int64).reshape(-1)
    bf = np.asarray(buffers, dtype=np.int64).reshape(-1)
    pi64 = ctypes.POINTER(ctypes.c_int64)
    n = ctypes.c_int64()
    args = (num_rows, len(nodes), nd.ctypes.data_as(pi64), len(buffers), bf.ctypes.data_as(pi64), body_length)
    L.check(L.lib().gz_ipc_record_batch_message(*args, None, 0, ctypes.byref(n)))
    out = np.zeros(n.value, np.uint8)
    L.check(L.lib().gz_ipc_record_batch_message(*args, out.ctypes.data_as(ctypes.c_void_p), n.value,
                                                ctypes.byref(n)))
    return out.tobytes()
