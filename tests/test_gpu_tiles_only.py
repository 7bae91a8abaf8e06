"""A matrix resident as tiles only (A's and A^T's blocked layouts built from
column ranges of A and row chunks of the host CSR, no CSR copies on the
device), the form the 100 GB matrix takes: its reconstructions are bitwise
those of the resident-CSR tiled path with the same tile shape (the tiles are
the same bytes), its sensitivity is bitwise the resident transpose's FP64
column sums, and within 1e-5 rel-L2 of the FP64 oracle."""
import numpy as np
import pytest

import mlem_amd as M
import oracle as O

pytestmark = pytest.mark.gpu

ITERS = 20


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def recon(m, g, f0, iters, **kw):
    o = M.opts(spmv=0, using_trans=0, **kw)
    with M.Recon(m, g, f0, options=o) as r:
        r.iterate(iters)
        r.sync()
        return r.image(), r.loglik()


@pytest.mark.parametrize("width,rows,chunk", [(1024, 512, 64 << 20), (1024, 2048, 0), (2048, 512, 24 << 20)])
def test_tiles_only_equals_resident_tiles(c1, width, rows, chunk):
    A, AT, g = c1["A"], c1["AT"], c1["g"]
    print(f"tiles only W={width} R={rows} chunk={chunk}", flush=True)
    f0 = np.ones(A.cols, np.float32)
    with M.ResidentMatrix(A, AT) as m:
        s32, s64 = m.sensitivity()
        ref_img, ref_ll = recon(m, g, f0, ITERS, tile_width=width, tile_rows=rows)
    with M.ResidentMatrix.tiles_only(A, tile_width=width, tile_rows=rows, chunk_bytes=chunk) as t:
        t32, t64 = t.sensitivity()
        img, ll = recon(t, g, f0, ITERS)
        assert t.info()["device_bytes"] > 0
    assert np.array_equal(s32.view(np.uint32), t32.view(np.uint32))
    assert np.array_equal(s64.view(np.uint64), t64.view(np.uint64))
    assert np.array_equal(img.view(np.uint32), ref_img.view(np.uint32))
    assert abs(ll - ref_ll) <= 1e-6 * abs(ref_ll)
    ref = O.mlem_f64(A.rowptr, A.col, A.val, A.rows, A.cols, g, s32.astype(np.float64), f0, ITERS)
    assert rel_l2(img, ref) <= 1e-5


@pytest.mark.parametrize("values", [1, 2, 3])
def test_tiles_only_16bit_values_equal_resident_tiles(c1, values):
    """16-bit values: A's codes quantized on the device before the column
    ranges (A^T's tiles hold them, the sensitivity is the quantized matrix's
    FP64 column sums), A's tiles quantized again from the host chunks: the
    images are bitwise those of the resident path with the same values."""
    A, AT, g = c1["A"], c1["AT"], c1["g"]
    f0 = np.ones(A.cols, np.float32)
    with M.ResidentMatrix(A, AT) as m:
        ref_img, _ = recon(m, g, f0, ITERS, tile_width=1024, tile_rows=512, values=values)
    with M.ResidentMatrix.tiles_only(A, tile_width=1024, tile_rows=512, values=values, chunk_bytes=48 << 20) as t:
        img, _ = recon(t, g, f0, ITERS)
        img2, _ = recon(t, g, f0, ITERS, values=values)
        with pytest.raises(M.MlemError):
            recon(t, g, f0, 1, values=(values % 3) + 1)  # not the matrix's values
    assert np.array_equal(img.view(np.uint32), ref_img.view(np.uint32))
    assert np.array_equal(img2.view(np.uint32), ref_img.view(np.uint32))


@pytest.mark.parametrize("per_pass,values,rows", [(4, 3, 2048), (4, 0, 512), (8, 2, 1024)])
def test_tiles_only_batches_equal_resident_batches(c1, per_pass, values, rows):
    """Tiles-only matrices in the batches' layout (quad groups, block
    streams) built from column ranges and row chunks: every slot of a batch
    bitwise the same as on the resident matrix; single reconstructions and
    other batch shapes are refused."""
    A, AT = c1["A"], c1["AT"]
    rng = np.random.default_rng(3)
    sinos = [rng.integers(0, 40, A.rows).astype(np.int32) for _ in range(per_pass + 2)]
    f0 = np.ones(A.cols, np.float32)

    def run(m, **kw):
        with M.Batch(m, len(sinos), values=values, per_pass=per_pass, tile_rows=rows, **kw) as b:
            for k, g in enumerate(sinos):
                b.set(k, g, f0)
            b.iterate(ITERS)
            return [b.image(k) for k in range(len(sinos))]

    with M.ResidentMatrix(A, AT) as m:
        ref = run(m)
    # scratch for a few column ranges (a range holds whole row blocks of A^T)
    with M.ResidentMatrix.tiles_only(A, tile_rows=rows, values=values, per_pass=per_pass,
                                     chunk_bytes=(48 << 20) * max(1, rows // 512)) as t:
        got = run(t)
        with pytest.raises(M.MlemError):
            recon(t, sinos[0], f0, 1)
        with pytest.raises(M.MlemError):
            M.Batch(t, 4, values=values, per_pass=1 if per_pass > 1 else 4)
    for k in range(len(sinos)):
        assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), k


def test_tiles_only_edge_matrices():
    """Empty rows and columns, one 60,000-entry row, heavy tails, duplicate
    columns, no entries, one row, row gaps: bitwise the resident tiled path,
    with scratch small enough for many column ranges and row chunks."""
    from matrices import make
    for kind in ["empty_rows", "one_huge_row", "heavy", "dups", "all_empty", "single_row", "row_gaps"]:
        A = make(kind, seed=3)
        AT = M.transpose(A)
        g = np.random.default_rng(4).integers(0, 30, A.rows).astype(np.int32)
        f0 = np.ones(A.cols, np.float32)
        print(f"tiles only edge {kind}: {A.rows}x{A.cols} nnz={A.nnz}", flush=True)
        with M.ResidentMatrix(A, AT) as m:
            ref_img, _ = recon(m, g, f0, 6, tile_width=1024, tile_rows=512)
        with M.ResidentMatrix.tiles_only(A, tile_width=1024, tile_rows=512, chunk_bytes=16 << 20) as t:
            img, _ = recon(t, g, f0, 6)
        assert np.array_equal(img.view(np.uint32), ref_img.view(np.uint32)), kind


def test_tiles_only_refuses_other_paths(c1):
    A, g = c1["A"], c1["g"]
    f0 = np.ones(A.cols, np.float32)
    with M.ResidentMatrix.tiles_only(A, tile_width=1024, tile_rows=512) as t:
        for kw in (dict(spmv=1), dict(spmv=0, using_trans=1), dict(spmv=0, tile_width=2048),
                   dict(spmv=0, values=1)):
            with pytest.raises(M.MlemError):
                M.Recon(t, g, f0, options=M.opts(**kw))
        with pytest.raises(M.MlemError):
            M.Batch(t, 4)
        with pytest.raises(M.MlemError):
            t.build_transpose()
        with pytest.raises(M.MlemError):
            t.autotune()
        with pytest.raises(M.MlemError):
            t.download()  # no CSR arrays on the device
    with pytest.raises(M.MlemError):
        M.ResidentMatrix.tiles_only(A, tile_width=512)
