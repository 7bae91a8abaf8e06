"""Seeded test matrices: the edge cases the reference's kernels meet
(empty rows, one huge row, heavy-tailed rows, duplicates within a row, no
nonzeros at all) as host CSR objects."""
import numpy as np

import mlem_amd as M


def from_lengths(lens, cols, seed, dup=False):
    rng = np.random.default_rng(seed)
    lens = np.asarray(lens, dtype=np.int64)
    rowptr = np.zeros(len(lens) + 1, np.uint64)
    rowptr[1:] = np.cumsum(lens)
    nnz = int(rowptr[-1])
    col = np.empty(nnz, np.uint32)
    for r in range(len(lens)):
        b, e = int(rowptr[r]), int(rowptr[r + 1])
        n = e - b
        if n == 0:
            continue
        if dup or n > cols:
            c = rng.integers(0, cols, n)
        else:
            c = rng.choice(cols, n, replace=False)
        col[b:e] = np.sort(c)
    val = (rng.random(nnz) * 2.0 + 0.01).astype(np.float32)
    return M.Csr(len(lens), cols, rowptr, col, val)


def make(kind, seed=0):
    rng = np.random.default_rng(seed + 1000)
    if kind == "ragged":
        return from_lengths(rng.integers(0, 70, 3000), 2000, seed)
    if kind == "heavy":
        lens = np.minimum((rng.pareto(1.1, 1500) * 40).astype(np.int64), 20000)
        return from_lengths(lens, 30000, seed)
    if kind == "empty_rows":
        lens = rng.integers(0, 40, 4000) * (rng.random(4000) < 0.3)
        return from_lengths(lens, 1000, seed)
    if kind == "one_huge_row":
        lens = np.concatenate([rng.integers(0, 5, 500), [60000], rng.integers(0, 5, 500)])
        return from_lengths(lens, 70000, seed)
    if kind == "all_empty":
        return from_lengths(np.zeros(777, np.int64), 50, seed)
    if kind == "single_row":
        return from_lengths([513], 600, seed)
    if kind == "dups":
        return from_lengths(rng.integers(1, 30, 800), 20, seed, dup=True)
    if kind == "tiny":
        return from_lengths([2, 0, 3], 4, seed)
    if kind == "row_gaps":
        # Non-empty rows separated by 0..12 empty ones (row steps 1..13 inside
        # a tile), rows of 1..20 entries over 5000 columns (several segments).
        gaps = rng.integers(0, 13, 3000)
        lens = np.zeros(int(gaps.sum()) + 3000, np.int64)
        pos = np.cumsum(gaps + 1) - 1
        lens[pos] = rng.integers(1, 21, 3000)
        return from_lengths(lens, 5000, seed)
    raise ValueError(kind)


KINDS = ["ragged", "heavy", "empty_rows", "one_huge_row", "all_empty", "single_row", "dups", "tiny", "row_gaps"]
