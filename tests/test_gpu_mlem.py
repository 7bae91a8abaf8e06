"""Full MLEM on the GPU against the FP64 oracle (oracle_mlem_f64).

Bar: relative L2 error of the FP32 image after N iterations <= 1e-5 against
the FP64 oracle run from the same start image with the same sensitivity.
"""
import numpy as np
import pytest

import mlem_amd as M
import oracle as O
from matrices import make

pytestmark = pytest.mark.gpu

REL_L2 = 1e-5


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)


@pytest.fixture(scope="module")
def resident(c1):
    m = M.ResidentMatrix(c1["A"], c1["AT"])
    yield m
    m.close()


@pytest.fixture(scope="module")
def ref50(c1, resident):
    A = c1["A"]
    s32, s64 = resident.sensitivity()
    f0 = np.ones(A.cols, np.float32)  # uniform x0 = 1
    ref = O.mlem_f64(A.rowptr, A.col, A.val, A.rows, A.cols, c1["g"], s32.astype(np.float64), f0, 50)
    return dict(s32=s32, s64=s64, f0=f0, ref=ref)


@pytest.mark.parametrize("trans,spmv", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_mlem_50_iterations_vs_fp64_oracle(c1, resident, ref50, trans, spmv):
    print(f"mlem C1 trans={trans} spmv={spmv}", flush=True)
    o = M.opts(spmv=spmv, using_trans=trans, secsize_fp=5, secsize_bp=5)
    img = M.mlem_run(resident, c1["g"], ref50["f0"], 50, sens=ref50["s32"], options=o)
    err = rel_l2(img, ref50["ref"])
    print(f"  rel-L2 vs FP64 oracle: {err:.3e}", flush=True)
    assert err <= REL_L2


@pytest.mark.parametrize("tile", [0, 1])
def test_merge_fp_with_csr_vector_bp_vs_fp64_oracle(c1, resident, ref50, tile):
    """The 2D workload's own pairing: merge-based FP (blocked, or the plain merge path) with CSR-vector BP on A^T
    (bp_spmv = 2), 50 iterations from x0 = 1."""
    o = M.opts(spmv=0, using_trans=0, secsize_fp=5, secsize_bp=5, tile_width=tile, bp_spmv=2)
    img = M.mlem_run(resident, c1["g"], ref50["f0"], 50, sens=ref50["s32"], options=o)
    err = rel_l2(img, ref50["ref"])
    print(f"merge FP (tile {tile}) + CSR-vector BP: rel-L2 vs FP64 oracle {err:.3e}", flush=True)
    assert err <= REL_L2


def test_sensitivity_fp64_matches_oracle(c1, resident, ref50):
    A = c1["A"]
    ref = O.colsum_f64(A.rowptr, A.col, A.val, A.rows, A.cols)
    np.testing.assert_allclose(ref50["s64"], ref, rtol=1e-12, atol=0)
    assert np.array_equal(ref50["s32"], ref.astype(np.float32))


def test_gpu_sensitivity_default_path(c1, resident, ref50):
    # sens=None: the reconstruction computes FP64 column sums itself.
    o = M.opts(spmv=1, using_trans=0)
    a = M.mlem_run(resident, c1["g"], ref50["f0"], 5, sens=None, options=o)
    b = M.mlem_run(resident, c1["g"], ref50["f0"], 5, sens=ref50["s32"], options=o)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_reference_init_and_scale_invariance(c1, resident, ref50):
    A, g = c1["A"], c1["g"]
    f0, init = M.initial_image(g, A.rows, ref50["s32"])
    assert init == O.initial_image(g, A.rows, ref50["s32"])
    o = M.opts(spmv=0, using_trans=0)
    img = M.mlem_run(resident, g, f0, 50, sens=ref50["s32"], options=o)
    # After >= 1 iteration MLEM does not depend on a uniform x0 scale.
    assert rel_l2(img, ref50["ref"]) <= REL_L2


def test_deterministic_and_graph_replay(c1, resident, ref50):
    o = M.opts(spmv=0, using_trans=0, secsize_fp=8, secsize_bp=8)
    a = M.mlem_run(resident, c1["g"], ref50["f0"], 20, sens=ref50["s32"], options=o)
    b = M.mlem_run(resident, c1["g"], ref50["f0"], 20, sens=ref50["s32"], options=o)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    og = M.opts(spmv=0, using_trans=0, secsize_fp=8, secsize_bp=8, use_graph=1)
    with M.Recon(resident, c1["g"], ref50["f0"], sens=ref50["s32"], options=og) as r:
        for _ in range(4):
            r.iterate(5)
            r.sync()
            r.image()  # eager host work between replays
        c = r.image()
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32))


def test_recon_loglik_fp64(c1, resident, ref50):
    A, g = c1["A"], c1["g"]
    o = M.opts(spmv=1, using_trans=0)
    with M.Recon(resident, g, ref50["f0"], sens=ref50["s32"], options=o) as r:
        lls = []
        for _ in range(6):
            lls.append(r.loglik())
            img = r.image()
            ref = O.loglik_f64(A.rowptr, A.col, A.val, A.rows, g, img.astype(np.float64))
            assert abs(lls[-1] - ref) <= 1e-6 * abs(ref)
            r.iterate(3)
    assert all(b >= a - 1e-9 * abs(a) for a, b in zip(lls, lls[1:]))  # EM ascent


@pytest.mark.parametrize("tile", [0, 1])
def test_profile_times(c1, resident, ref50, tile):
    o = M.opts(spmv=0, using_trans=0, profile=1, tile_width=tile)
    with M.Recon(resident, c1["g"], ref50["f0"], sens=ref50["s32"], options=o) as r:
        r.iterate(10)
        r.sync()
        fp, bp, n = r.kernel_times()
        ph, n2 = r.phase_times()
    assert n == 10 and n2 == 10 and fp > 0 and bp > 0
    assert all(v >= 0 for v in ph) and ph[0] > 0 and ph[2] > 0
    assert abs(ph[0] + ph[1] - fp) <= 1e-9 * fp and abs(ph[2] + ph[3] - bp) <= 1e-9 * bp


def test_copy_kernel():
    n = (1 << 20) + 3
    a = M.DeviceArray(n, np.float32)
    b = M.DeviceArray(n, np.float32)
    x = np.random.default_rng(0).random(n).astype(np.float32)
    a.upload(x)
    M.copy_f32(b, a)
    assert np.array_equal(b.numpy(), x)


def test_edge_matrices_full_mlem():
    for kind in ["empty_rows", "one_huge_row", "heavy", "dups"]:
        A = make(kind, seed=5)
        AT = M.transpose(A)
        rng = np.random.default_rng(4)
        g = rng.integers(0, 30, A.rows).astype(np.int32)
        with M.ResidentMatrix(A, AT) as m:
            s32, _ = m.sensitivity()
            f0 = np.ones(A.cols, np.float32)
            ref = O.mlem_f64(A.rowptr, A.col, A.val, A.rows, A.cols, g, s32.astype(np.float64), f0, 10)
            for trans, spmv in [(0, 0), (0, 1), (1, 1)]:
                print(f"edge {kind} trans={trans} spmv={spmv}", flush=True)
                img = M.mlem_run(m, g, f0, 10, sens=s32, options=M.opts(spmv=spmv, using_trans=trans, secsize_fp=3,
                                                                         secsize_bp=4))
                assert rel_l2(img, ref) <= (REL_L2 if trans == 0 else 1e-4)


def test_using_trans_requires_transpose(c1):
    with M.ResidentMatrix(c1["A"]) as m:
        with pytest.raises(M.MlemError):
            M.Recon(m, c1["g"], np.ones(c1["A"].cols, np.float32), options=M.opts(using_trans=0))
        # scatter path works without a transpose
        img = M.mlem_run(m, c1["g"], np.ones(c1["A"].cols, np.float32), 2, options=M.opts(using_trans=1, spmv=1))
        assert np.isfinite(img).all()


@pytest.mark.parametrize("width,rows", [(1024, 512), (1024, 1024), (1024, 2048), (1024, 4096), (1024, 16384),
                                        (2048, 512), (2048, 2048), (2048, 8192)])
def test_blocked_merge_path_vs_fp64_oracle(c1, resident, ref50, width, rows):
    """Merge-based FP and BP through the LDS-blocked kernels, both segment widths and several block sizes."""
    for secsize in (8, 16):
        print(f"blocked W={width} R={rows} S={secsize}", flush=True)
        o = M.opts(spmv=0, using_trans=0, secsize_fp=secsize, secsize_bp=secsize, tile_width=width, tile_rows=rows)
        img = M.mlem_run(resident, c1["g"], ref50["f0"], 50, sens=ref50["s32"], options=o)
        assert rel_l2(img, ref50["ref"]) <= REL_L2
        img2 = M.mlem_run(resident, c1["g"], ref50["f0"], 50, sens=ref50["s32"], options=o)
        assert np.array_equal(img.view(np.uint32), img2.view(np.uint32))


def test_blocked_edge_matrices():
    for kind in ["empty_rows", "one_huge_row", "heavy", "dups", "all_empty", "single_row", "row_gaps"]:
        A = make(kind, seed=6)
        AT = M.transpose(A)
        rng = np.random.default_rng(5)
        g = rng.integers(0, 30, A.rows).astype(np.int32)
        with M.ResidentMatrix(A, AT) as m:
            s32, _ = m.sensitivity()
            f0 = np.ones(A.cols, np.float32)
            ref = O.mlem_f64(A.rowptr, A.col, A.val, A.rows, A.cols, g, s32.astype(np.float64), f0, 8)
            for width, rows in ((1024, 512), (2048, 4096)):
                print(f"blocked edge {kind} W={width} R={rows}", flush=True)
                o = M.opts(spmv=0, using_trans=0, secsize_fp=8, secsize_bp=16, tile_width=width, tile_rows=rows)
                img = M.mlem_run(m, g, f0, 8, sens=s32, options=o)
                assert rel_l2(img, ref) <= REL_L2 or np.linalg.norm(ref) == 0


@pytest.mark.parametrize("groups", [1, 7, 64, 100000])
def test_blocked_work_split(c1, resident, ref50, groups):
    """Any number of runs per row block (one run per block up to one unit per run) gives the same image."""
    print(f"blocked groups={groups}", flush=True)
    o = M.opts(spmv=0, secsize_fp=8, secsize_bp=16, tile_width=1024, tile_rows=512, tile_groups=groups)
    img = M.mlem_run(resident, c1["g"], ref50["f0"], 50, sens=ref50["s32"], options=o)
    assert rel_l2(img, ref50["ref"]) <= REL_L2


def test_blocked_column_segment_order_is_bitwise_row_order(c1, ref50, monkeypatch):
    """MLEM_TILE_ORDER=cols launches the blocked kernel's runs by column segment instead of by row block: each
    run still writes its own partial sums and the epilogue adds them in run order, so the images are bitwise the
    row-order ones (two tile shapes, several workgroup targets, 30 iterations from x0 = 1)."""
    A, AT, g, f0, s32 = c1["A"], c1["AT"], c1["g"], ref50["f0"], ref50["s32"]
    for kw in (dict(tile_width=1024, tile_rows=512), dict(tile_width=2048, tile_rows=2048, tile_groups=64),
               dict(tile_width=1024, tile_rows=1024, tile_groups=100000)):
        imgs = []
        for order in ("rows", "cols"):
            monkeypatch.setenv("MLEM_TILE_ORDER", order)
            print(f"blocked launch order {order} {kw}", flush=True)
            with M.ResidentMatrix(A, AT) as m:  # tiles and work lists are built per matrix, under the order set
                imgs.append(M.mlem_run(m, g, f0, 30, sens=s32, options=M.opts(spmv=0, using_trans=0, **kw)))
        assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
        assert np.isfinite(imgs[0]).all() and imgs[0].sum() > 0


def test_default_row_blocks_follow_matrix_size(c1, resident, ref50):
    """The 2D ring has under two units of entries per run at the default workgroup counts, so its default row
    blocks are 512 rows: the default run is bitwise the 512-row run (one reconstruction and a batch of 4 per pass),
    not the 2048-row one, and both are within the oracle bar.  A small workgroup target (runs of about four units)
    gives the same matrix 2048-row blocks again."""
    g, f0, s32 = c1["g"], ref50["f0"], ref50["s32"]
    run = lambda **kw: M.mlem_run(resident, g, f0, 50, sens=s32, options=M.opts(spmv=0, using_trans=0, **kw))
    auto, r512, r2048 = run(), run(tile_rows=512), run(tile_rows=2048)
    assert np.array_equal(auto.view(np.uint32), r512.view(np.uint32))
    assert not np.array_equal(auto.view(np.uint32), r2048.view(np.uint32))
    for img in (auto, r2048):
        assert rel_l2(img, ref50["ref"]) <= REL_L2
    # runs of about four units of 2048 entries each
    groups = int(c1["A"].nnz // 4096 // 2)
    few_auto, few_2048 = run(tile_groups=groups), run(tile_groups=groups, tile_rows=2048)
    assert np.array_equal(few_auto.view(np.uint32), few_2048.view(np.uint32))
    imgs = []
    for rows in (0, 512):
        with M.Batch(resident, 4, per_pass=4, tile_rows=rows) as b:
            for k in range(4):
                b.set(k, g, f0)
            b.iterate(10)
            imgs.append(b.image(2))
    assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))


def test_blocked_largest_tiles(c1, resident, ref50):
    # 16384 rows x 2048 columns per tile: 72 KiB of LDS per workgroup.
    o = M.opts(spmv=0, tile_width=2048, tile_rows=16384, secsize_fp=16, secsize_bp=16)
    img = M.mlem_run(resident, c1["g"], ref50["f0"], 50, sens=ref50["s32"], options=o)
    assert rel_l2(img, ref50["ref"]) <= REL_L2


def test_tiling_refused_and_automatic_fallback():
    # Unsorted rows: tiles are refused; automatic mode falls back to the
    # plain merge path, an explicit width is an error.
    A = make("ragged", seed=2)
    col = A.col.copy()
    for r in range(0, A.rows, 3):
        b, e = int(A.rowptr[r]), int(A.rowptr[r + 1])
        col[b:e] = col[b:e][::-1]
    U = M.Csr(A.rows, A.cols, A.rowptr, col, A.val)
    g = np.random.default_rng(1).integers(0, 20, A.rows).astype(np.int32)
    f0 = np.ones(A.cols, np.float32)
    with M.ResidentMatrix(U, M.transpose(U)) as m:
        s32, _ = m.sensitivity()
        ref = O.mlem_f64(U.rowptr, U.col, U.val, U.rows, U.cols, g, s32.astype(np.float64), f0, 5)
        img = M.mlem_run(m, g, f0, 5, sens=s32, options=M.opts(spmv=0, tile_width=0))
        assert rel_l2(img, ref) <= REL_L2
        with pytest.raises(M.MlemError):
            M.mlem_run(m, g, f0, 5, sens=s32, options=M.opts(spmv=0, bp_spmv=2, tile_width=2048))
    # A dense block (tiles of more than 65536 entries) is fine.
    D = __import__("matrices").from_lengths(np.full(1024, 700), 2048, seed=3)
    g = np.random.default_rng(2).integers(1, 9, D.rows).astype(np.int32)
    with M.ResidentMatrix(D, M.transpose(D)) as m:
        s32, _ = m.sensitivity()
        f0 = np.ones(D.cols, np.float32)
        ref = O.mlem_f64(D.rowptr, D.col, D.val, D.rows, D.cols, g, s32.astype(np.float64), f0, 3)
        img = M.mlem_run(m, g, f0, 3, sens=s32, options=M.opts(spmv=0, tile_width=2048, tile_rows=1024))
        assert rel_l2(img, ref) <= REL_L2


def test_matrix_refuses_transpose_rebuild_and_destroy_while_in_use(c1):
    """A reconstruction or batch holds the matrix's transpose, tiles and graphs by address: rebuilding the
    transpose or destroying the matrix is refused while one is open, and allowed again once they are closed."""
    m = M.ResidentMatrix(c1["A"])
    m.build_transpose()
    f0 = np.ones(c1["A"].cols, np.float32)
    r = M.Recon(m, c1["g"], f0, options=M.opts(spmv=0, using_trans=0))
    b = M.Batch(m, 2, per_pass=4)
    for closer in (r, b):
        with pytest.raises(M.MlemError, match="still use"):
            m.build_transpose()
        with pytest.raises(M.MlemError, match="still use"):
            m.close()
        assert m.h
        closer.close()
    m.build_transpose()
    m.close()
    assert not m.h
