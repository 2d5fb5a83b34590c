"""Raw PhaseField / grain files in the reference's binary formats.

PhaseField raw .dat (PhaseField::Write/Read, src/PhaseField.cpp:2340-2388, 2487-2543; NodePF::write/
read, include/Containers/NodePF.h:751-775): three native-endian int32 Nx, Ny, Nz (storage sizes,
reduced axes written as 1), then per cell in x-outer, z-inner order a size_t entry count followed
by that many (size_t index, double value) pairs.
"""
from __future__ import annotations

import numpy as np

from .microstructure import Lists


def write_phasefield(path: str, lists: Lists) -> None:
    n = lists.ncells
    off = lists.offsets()
    # one record per cell: count (u64) + count x (u64 index, f64 value)
    rec = np.empty(n + 2 * int(lists.count.sum()), dtype=np.uint64)
    pos = np.arange(n, dtype=np.int64) + 2 * off[:-1]
    rec[pos] = lists.count.astype(np.uint64)
    e = np.arange(int(lists.count.sum()), dtype=np.int64)
    cell_of = np.repeat(np.arange(n, dtype=np.int64), lists.count)
    epos = pos[cell_of] + 1 + 2 * (e - off[cell_of])
    rec[epos] = lists.index.astype(np.uint64)
    rec[epos + 1] = lists.value.astype(np.float64).view(np.uint64)
    with open(path, "wb") as f:
        f.write(np.asarray(lists.shape, dtype=np.int32).tobytes())
        f.write(rec.tobytes())


def read_phasefield(path: str) -> Lists:
    raw = np.fromfile(path, dtype=np.uint8)
    shape = tuple(int(v) for v in raw[:12].view(np.int32))
    body = raw[12:].view(np.uint64)
    n = shape[0] * shape[1] * shape[2]
    count = np.empty(n, dtype=np.int32)
    pos = 0
    # walk the records (counts are small); vectorised per pass over cells is not possible
    # without knowing the counts, so do it in one Python loop over cells in chunks
    idx_list = []
    val_list = []
    for c in range(n):
        k = int(body[pos])
        count[c] = k
        if k:
            blk = body[pos + 1: pos + 1 + 2 * k]
            idx_list.append(blk[0::2].astype(np.int64))
            val_list.append(blk[1::2].view(np.float64).copy())
        pos += 1 + 2 * k
    index = np.concatenate(idx_list).astype(np.int32) if idx_list else np.zeros(0, np.int32)
    value = np.concatenate(val_list) if val_list else np.zeros(0, np.float64)
    return Lists(shape, count, index, value)


def write_grains(path: str, phase, quaternions=None) -> None:
    """Grain table for the golden-vector driver: uint32 n, then per grain uint32 phase and four
    doubles (s, x, y, z) of the orientation quaternion."""
    phase = np.asarray(phase, dtype=np.uint32)
    n = phase.shape[0]
    q = np.zeros((n, 4)) if quaternions is None else np.asarray(quaternions, dtype=np.float64)
    rec = np.zeros(n, dtype=[("p", "<u4"), ("q", "<f8", (4,))])
    rec["p"] = phase
    rec["q"] = q
    with open(path, "wb") as f:
        f.write(np.uint32(n).tobytes())
        f.write(rec.tobytes())
