"""mlem_lor_order: the row order of the batched layouts, from each LOR's
transaxial line fitted to its voxels.

Pinned by what the geometry fixes: on a hand-made grid, rows that are exact
lines at known angles and offsets come out sorted by (10-degree angle block,
offset, angle); on the scanner matrices the order is a permutation, the same
for any thread count, and it lowers the quad padding of A^T's 1024-wide
column segments (the quantity it exists for), counted here on the host."""
import numpy as np
import pytest

import mlem_amd as M


def line_rows(nx, ny, lines):
    """A CSR with one row per (angle in degrees, offset in voxels): the voxels within 0.5 of the line
    x cos(phi + 90) + y sin(phi + 90) = offset through the grid centre, weight 1."""
    cx, cy = (nx - 1) / 2, (ny - 1) / 2
    yy, xx = np.mgrid[0:ny, 0:nx]
    rowptr, cols = [0], []
    for deg, off in lines:
        t = np.deg2rad(deg)
        # signed distance from the line whose direction is (cos t, sin t), offset `off` along its normal
        d = -(xx - cx) * np.sin(t) + (yy - cy) * np.cos(t) - off
        sel = np.flatnonzero(np.abs(d).ravel() <= 0.5)
        cols.append(sel)
        rowptr.append(rowptr[-1] + len(sel))
    col = np.concatenate(cols).astype(np.uint32)
    return M.Csr(len(lines), nx * ny, np.array(rowptr, np.uint64), col, np.ones(len(col), np.float32))


def test_lines_sort_by_angle_block_then_offset():
    nx = ny = 64
    # angles in different 10-degree blocks, offsets mixed; the expected order is (block, offset)
    lines = [(47.0, 3.0), (2.0, -5.0), (95.0, 0.0), (2.0, 7.0), (47.0, -9.0), (133.0, 1.0), (2.0, 0.0),
             (95.0, -4.0), (170.0, 2.0), (47.0, 0.0)]
    A = line_rows(nx, ny, lines)
    order = M.lor_order(A, (nx, ny, 1))
    placed = [lines[i] for i in np.argsort(order)]
    assert [int(d // 10) for d, _ in placed] == sorted(int(d // 10) for d, _ in lines)
    for blk in {int(d // 10) for d, _ in lines}:
        offs = [o for d, o in placed if int(d // 10) == blk]
        assert offs == sorted(offs), (blk, offs)


def test_empty_rows_go_last_in_row_order():
    nx = ny = 16
    A = line_rows(nx, ny, [(30.0, 0.0), (120.0, 2.0)])
    # insert two empty rows: rows 0 and 2
    rowptr = np.array([0, 0, A.rowptr[1], A.rowptr[1], A.rowptr[2]], np.uint64)
    B = M.Csr(4, A.cols, rowptr, A.col, A.val)
    order = M.lor_order(B, (nx, ny, 1))
    assert sorted(order.tolist()) == [0, 1, 2, 3]
    assert order[0] == 2 and order[2] == 3


def quad_padding(AT, perm, W=1024, sample=3000, seed=0):
    rng = np.random.default_rng(seed)
    rows = np.sort(rng.choice(AT.rows, min(AT.rows, sample), replace=False))
    tp = AT.rowptr.astype(np.int64)
    stored = nn = 0
    for r in rows:
        c = AT.col[tp[r]:tp[r + 1]].astype(np.int64)
        if len(c) == 0:
            continue
        if perm is not None:
            c = perm[c]
        _, cnt = np.unique(c // W, return_counts=True)
        stored += int(((cnt + 3) // 4 * 4).sum())
        nn += len(c)
    return stored / nn - 1.0


@pytest.mark.parametrize("geo_name", ["RING2D_288", "RING3D_576_MID"])
def test_order_is_a_deterministic_permutation_that_cuts_padding(geo_name):
    geo = getattr(M, geo_name)()
    A = M.gen_system_matrix(geo, seed=1)
    grid = tuple(geo.grid)
    order = M.lor_order(A, grid)
    assert np.array_equal(np.sort(order), np.arange(A.rows, dtype=np.uint32))
    assert np.array_equal(order, M.lor_order(A, grid, threads=1))
    assert np.array_equal(order, M.lor_order(A, grid, threads=3))
    AT = M.transpose(A, values=False)
    before, after = quad_padding(AT, None), quad_padding(AT, order.astype(np.int64))
    print(f"{geo_name}: A^T quad padding at W=1024 {100 * before:.1f}% -> {100 * after:.1f}%")
    assert after < 0.75 * before


def test_bad_arguments():
    A = line_rows(8, 8, [(10.0, 0.0)])
    with pytest.raises(M.MlemError):
        M.lor_order(A, (8, 4, 1))  # 32 voxels, not 64
    bad = M.Csr(1, 64, A.rowptr, A.col.copy(), A.val)
    bad.col[0] = 64
    with pytest.raises(M.MlemError):
        M.lor_order(bad, (8, 8, 1))
