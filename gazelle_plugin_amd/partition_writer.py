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
