"""Kernel-level parity on the GPU: every entry point of the kernel C ABI
against the FP64 oracle (oracle/), on the 2D ring matrix and on edge-case
matrices.

Tolerances (FP32 kernels vs the FP64 oracle, all inputs non-negative):
  * SpMV / fused outputs: |y - y64| <= 2e-5 * |y64| + 1e-30 per element
    (sums of non-negative terms: the error is bounded by the longest serial
    chain of float additions times 2^-24, below 2e-5 for these shapes);
  * elementwise ratio / update: bit-exact with the same FP32 formula;
  * FP64 sensitivity: 1e-12 relative;
  * transpose: bit-exact.
"""
import numpy as np
import pytest

import mlem_amd as M
import oracle as O
from matrices import KINDS, make

pytestmark = pytest.mark.gpu

RTOL = 2e-5


def upload(A):
    return (M.DeviceArray.from_numpy(A.rowptr), M.DeviceArray.from_numpy(A.col), M.DeviceArray.from_numpy(A.val))


def close(y, ref, rtol=RTOL):
    y = np.asarray(y, np.float64)
    err = np.abs(y - ref)
    bad = err > rtol * np.abs(ref) + 1e-30
    assert not bad.any(), f"{bad.sum()} elements off, worst rel {np.max(err / (np.abs(ref) + 1e-30)):.3g}"


def matrices(c1):
    out = [("c1", c1["A"])]
    out += [(k, make(k, seed=3)) for k in KINDS]
    return out


@pytest.fixture(scope="module")
def mats(c1):
    return matrices(c1)


@pytest.mark.parametrize("secsize", [1, 3, 5, 8, 16, 32])
def test_fp_merge_matches_oracle(mats, secsize):
    for name, A in mats:
        print(f"fp_merge {name} S={secsize} rows={A.rows} nnz={A.nnz}", flush=True)
        rng = np.random.default_rng(7)
        x = (rng.random(A.cols) + 0.1).astype(np.float32)
        d_rp, d_c, d_v = upload(A)
        d_x = M.DeviceArray.from_numpy(x)
        d_y = M.DeviceArray(max(A.rows, 1), np.float32)
        wsb = M.merge_workspace_bytes(A.rows, A.nnz, secsize)
        ws = M.DeviceArray(wsb, np.uint8)
        M.fp_merge(d_rp, d_c, d_v, d_x, d_y, secsize, A.rows, A.nnz, ws, wsb)
        y = d_y.numpy()[:A.rows]
        ref = O.spmv_f64(A.rowptr, A.col, A.val, x.astype(np.float64), A.rows)
        close(y, ref)
        # Deterministic: a second call gives the same bits.
        M.fp_merge(d_rp, d_c, d_v, d_x, d_y, secsize, A.rows, A.nnz, ws, wsb)
        assert np.array_equal(d_y.numpy()[:A.rows].view(np.uint32), y.view(np.uint32))


@pytest.mark.parametrize("lanes", [0, 4, 8, 16, 32, 64])
def test_fp_csr_vector_matches_oracle(mats, lanes):
    for name, A in mats:
        print(f"fp_csr_vector {name} lanes={lanes}", flush=True)
        rng = np.random.default_rng(8)
        x = (rng.random(A.cols) + 0.1).astype(np.float32)
        d_rp, d_c, d_v = upload(A)
        d_x = M.DeviceArray.from_numpy(x)
        d_y = M.DeviceArray(max(A.rows, 1), np.float32)
        M.fp_csr_vector(d_rp, d_c, d_v, d_x, d_y, A.rows, A.nnz, lanes)
        y = d_y.numpy()[:A.rows]
        close(y, O.spmv_f64(A.rowptr, A.col, A.val, x.astype(np.float64), A.rows))
        M.fp_csr_vector(d_rp, d_c, d_v, d_x, d_y, A.rows, A.nnz, lanes)
        assert np.array_equal(d_y.numpy()[:A.rows].view(np.uint32), y.view(np.uint32))


def test_bp_on_transpose_matches_oracle(mats):
    for name, A in mats:
        print(f"bp {name}", flush=True)
        AT = M.transpose(A)
        rng = np.random.default_rng(9)
        r = (rng.random(A.rows) + 0.1).astype(np.float32)
        ref = O.spmv_f64(AT.rowptr, AT.col, AT.val, r.astype(np.float64), AT.rows)
        d_rp, d_c, d_v = upload(AT)
        d_r = M.DeviceArray.from_numpy(r)
        d_y = M.DeviceArray(max(AT.rows, 1), np.float32)
        M.bp_csr_vector(d_rp, d_c, d_v, d_r, d_y, AT.rows, AT.nnz)
        close(d_y.numpy()[:AT.rows], ref)
        for s in (2, 5, 9):
            wsb = M.merge_workspace_bytes(AT.rows, AT.nnz, s)
            ws = M.DeviceArray(wsb, np.uint8)
            M.bp_merge(d_rp, d_c, d_v, d_r, d_y, s, AT.rows, AT.nnz, ws, wsb)
            close(d_y.numpy()[:AT.rows], ref)


def test_bp_scatter_matches_oracle(mats):
    for name, A in mats:
        print(f"bp_scatter {name}", flush=True)
        rng = np.random.default_rng(10)
        r = (rng.random(A.rows) + 0.1).astype(np.float32)
        r[::7] = 0.0
        AT = M.transpose(A)
        ref = O.spmv_f64(AT.rowptr, AT.col, AT.val, r.astype(np.float64), AT.rows)
        d_rp, d_c, d_v = upload(A)
        d_r = M.DeviceArray.from_numpy(r)
        d_y = M.DeviceArray.from_numpy(np.full(A.cols, 123.0, np.float32))  # overwritten, not accumulated
        M.bp_scatter(d_rp, d_c, d_v, d_r, d_y, A.rows, A.cols, A.nnz)
        close(d_y.numpy(), ref, rtol=1e-4)


def test_fused_fp_ratio(mats):
    for name, A in mats:
        print(f"fp_ratio {name}", flush=True)
        rng = np.random.default_rng(11)
        x = (rng.random(A.cols) + 0.1).astype(np.float32)
        g = rng.integers(0, 50, A.rows).astype(np.int32)
        fp = O.spmv_f64(A.rowptr, A.col, A.val, x.astype(np.float64), A.rows)
        ref = np.where(fp != 0, g / np.where(fp != 0, fp, 1.0), 0.0)
        d_rp, d_c, d_v = upload(A)
        d_x, d_g = M.DeviceArray.from_numpy(x), M.DeviceArray.from_numpy(g)
        d_y = M.DeviceArray(max(A.rows, 1), np.float32)
        M.fp_ratio_csr_vector(d_rp, d_c, d_v, d_x, d_g, d_y, A.rows, A.nnz)
        close(d_y.numpy()[:A.rows], ref, rtol=4e-5)
        wsb = M.merge_workspace_bytes(A.rows, A.nnz, 6)
        ws = M.DeviceArray(wsb, np.uint8)
        M.fp_ratio_merge(d_rp, d_c, d_v, d_x, d_g, d_y, 6, A.rows, A.nnz, ws, wsb)
        close(d_y.numpy()[:A.rows], ref, rtol=4e-5)


def test_fused_bp_update(mats):
    for name, A in mats:
        print(f"bp_update {name}", flush=True)
        AT = M.transpose(A)
        rng = np.random.default_rng(12)
        r = (rng.random(A.rows) + 0.1).astype(np.float32)
        f = (rng.random(A.cols) + 0.5).astype(np.float32)
        sens = (rng.random(A.cols) + 0.2).astype(np.float32)
        sens[::5] = 0.0
        bp = O.spmv_f64(AT.rowptr, AT.col, AT.val, r.astype(np.float64), AT.rows)
        ref = np.where(sens == 0, f * bp, f * bp / np.where(sens == 0, 1, sens))
        d_rp, d_c, d_v = upload(AT)
        d_r, d_s = M.DeviceArray.from_numpy(r), M.DeviceArray.from_numpy(sens)
        d_f = M.DeviceArray.from_numpy(f)
        M.bp_update_csr_vector(d_rp, d_c, d_v, d_r, d_s, d_f, AT.rows, AT.nnz)
        close(d_f.numpy(), ref, rtol=4e-5)
        d_f.upload(f)
        wsb = M.merge_workspace_bytes(AT.rows, AT.nnz, 7)
        ws = M.DeviceArray(wsb, np.uint8)
        M.bp_update_merge(d_rp, d_c, d_v, d_r, d_s, d_f, 7, AT.rows, AT.nnz, ws, wsb)
        close(d_f.numpy(), ref, rtol=4e-5)


def test_elementwise_ratio_update_bit_exact():
    rng = np.random.default_rng(13)
    n = 100_003
    g = rng.integers(0, 1000, n).astype(np.int32)
    fw = (rng.random(n) * 5).astype(np.float32)
    fw[::9] = 0.0
    d_g, d_fw = M.DeviceArray.from_numpy(g), M.DeviceArray.from_numpy(fw)
    M.ratio(d_g, d_fw, n)
    with np.errstate(divide="ignore", invalid="ignore"):
        ref = np.where(fw != 0, g.astype(np.float32) / fw, fw).astype(np.float32)
    assert np.array_equal(d_fw.numpy().view(np.uint32), ref.view(np.uint32))
    f = (rng.random(n) + 0.5).astype(np.float32)
    nm = (rng.random(n) + 0.1).astype(np.float32)
    nm[::4] = 0
    bw = rng.random(n).astype(np.float32)
    d_f, d_nm, d_bw = M.DeviceArray.from_numpy(f), M.DeviceArray.from_numpy(nm), M.DeviceArray.from_numpy(bw)
    M.update(d_f, d_nm, d_bw, n)
    v = (f * bw).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        ref = np.where(nm == 0, v, v / nm).astype(np.float32)
    assert np.array_equal(d_bw.numpy().view(np.uint32), ref.view(np.uint32))


def test_colsum_f64(mats):
    for name, A in mats:
        print(f"colsum {name}", flush=True)
        AT = M.transpose(A)
        ref = O.colsum_f64(A.rowptr, A.col, A.val, A.rows, A.cols)
        d_rp, d_c, d_v = upload(AT)
        s64 = M.DeviceArray(A.cols, np.float64)
        s32 = M.DeviceArray(A.cols, np.float32)
        M.colsum_f64(d_rp, d_v, A.cols, A.nnz, s64, s32)
        np.testing.assert_allclose(s64.numpy(), ref, rtol=1e-12, atol=0)
        assert np.array_equal(s32.numpy(), ref.astype(np.float32))
        a_rp, a_c, a_v = upload(A)
        M.colsum_scatter_f64(a_rp, a_c, a_v, A.rows, A.cols, A.nnz, s64, s32)
        np.testing.assert_allclose(s64.numpy(), ref, rtol=1e-12, atol=0)


def test_device_transpose_bit_exact(mats):
    for name, A in mats:
        print(f"device transpose {name}", flush=True)
        host = M.transpose(A)
        m = M.ResidentMatrix(A)
        m.build_transpose()
        dev = m.download_transpose()
        assert np.array_equal(dev.rowptr, host.rowptr)
        assert np.array_equal(dev.col, host.col)
        assert np.array_equal(dev.val.view(np.uint32), host.val.view(np.uint32))
        cp, rt, vt = O.sptrans(A.rows, A.cols, A.rowptr, A.col, A.val, procs=3)
        assert np.array_equal(dev.rowptr, cp) and np.array_equal(dev.col, rt)
        m.close()


@pytest.mark.parametrize("budget", [2 << 20, 4 << 20, 64 << 20])
def test_device_transpose_in_column_ranges_bit_exact(mats, budget):
    """The transpose within a small scratch budget (column ranges sorted one
    at a time, as for the 100 GB matrices) is bit-identical to the host
    transpose; a budget too small for the fixed part is refused."""
    many = 0
    for name, A in mats:
        host = M.transpose(A)
        with M.ResidentMatrix(A) as m:
            ranges = m.build_transpose(budget)
            print(f"transpose in ranges {name} budget={budget} ranges={ranges}", flush=True)
            many = max(many, ranges)
            dev = m.download_transpose()
        assert np.array_equal(dev.rowptr, host.rowptr)
        assert np.array_equal(dev.col, host.col)
        assert np.array_equal(dev.val.view(np.uint32), host.val.view(np.uint32))
    if budget == 2 << 20:
        assert many > 1
    A = mats[0][1]
    with M.ResidentMatrix(A) as m:
        with pytest.raises(M.MlemError):
            m.build_transpose(4096)


def test_bad_arguments_rejected():
    A = make("tiny")
    d_rp, d_c, d_v = upload(A)
    d_x = M.DeviceArray(A.cols, np.float32)
    d_y = M.DeviceArray(A.rows, np.float32)
    with pytest.raises(M.MlemError):
        M.fp_merge(d_rp, d_c, d_v, d_x, d_y, 0, A.rows, A.nnz, None, 0)
    with pytest.raises(M.MlemError):
        M.fp_merge(d_rp, d_c, d_v, d_x, d_y, 4, A.rows, A.nnz, None, 0)  # no workspace
    with pytest.raises(M.MlemError):
        M.fp_csr_vector(d_rp, d_c, d_v, d_x, d_y, A.rows, A.nnz, lanes=3)
    bad = M.Csr(A.rows, A.cols, A.rowptr, A.col.copy(), A.val)
    bad.col[0] = A.cols  # out-of-range column must be caught on the host
    with pytest.raises(M.MlemError):
        M.ResidentMatrix(bad)
