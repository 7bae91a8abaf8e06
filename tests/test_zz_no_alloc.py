"""mlem_fp_merge_noalloc (the merge-path forward projection with the
library's fixed scratch instead of a workspace) through the C ABI: bitwise
the output of mlem_fp_merge with a workspace, on two calls in a row, the
second with a smaller input (the scratch keeps the first call's contents,
which must not leak into the second), and sizes beyond the fixed scratch
refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

import mlem_amd as M
from matrices import make

pytestmark = pytest.mark.gpu


def upload(A):
    return (M.DeviceArray.from_numpy(A.rowptr), M.DeviceArray.from_numpy(A.col), M.DeviceArray.from_numpy(A.val))


def both(A, x, secsize):
    """(mlem_fp_merge with a workspace, mlem_fp_merge_noalloc) outputs for y = A x."""
    d_rp, d_c, d_v = upload(A)
    d_x = M.DeviceArray.from_numpy(x)
    wsb = M.merge_workspace_bytes(A.rows, A.nnz, secsize)
    ws = M.DeviceArray(wsb, np.uint8)
    d_ref = M.DeviceArray(A.rows, np.float32)
    d_out = M.DeviceArray(A.rows, np.float32)
    M.fill_f32(d_out, 12345.0)  # overwritten, not accumulated
    M.fp_merge(d_rp, d_c, d_v, d_x, d_ref, secsize, A.rows, A.nnz, ws, wsb)
    M.fp_merge_noalloc(d_rp, d_c, d_v, d_x, d_out, secsize, A.rows, A.nnz)
    return d_ref.numpy(), d_out.numpy()


def test_noalloc_bitwise_the_workspace_entry_point_twice_second_smaller(c1):
    rng = np.random.default_rng(5)
    first = c1["A"]  # 41,328 rows, 2.0e6 nonzeros: 3,900 merge tiles at secsize 2
    second = make("heavy", seed=4)  # 1,500 rows, fewer tiles: stale scratch past its end must not matter
    assert second.rows + second.nnz < first.rows + first.nnz
    for A, secsize in ((first, 2), (second, 7)):
        print(f"noalloc fp_merge rows={A.rows} nnz={A.nnz} secsize={secsize}", flush=True)
        x = (rng.random(A.cols) + 0.1).astype(np.float32)
        ref, out = both(A, x, secsize)
        assert np.array_equal(ref.view(np.uint32), out.view(np.uint32))


def test_noalloc_refuses_sizes_beyond_its_scratch():
    A = make("tiny", seed=1)
    d_rp, d_c, d_v = upload(A)
    d_x = M.DeviceArray.from_numpy(np.ones(A.cols, np.float32))
    d_y = M.DeviceArray(A.rows, np.float32)
    # rows + nnz claimed far beyond what 32 MiB of partition and carries cover at secsize 1: refused on the host
    # before any kernel reads the (small) arrays
    big_nnz = 1 << 31
    assert M.merge_workspace_bytes(A.rows, big_nnz, 1) > 32 << 20
    with pytest.raises(M.MlemError, match="scratch"):
        M.fp_merge_noalloc(d_rp, d_c, d_v, d_x, d_y, 1, A.rows, big_nnz)
    lib = M.hip_lib()
    assert lib.mlem_fp_merge_noalloc(None, None, None, None, None, 1, 3, 5, None) == 1  # null pointers
    assert lib.mlem_fp_merge_noalloc(d_rp.h, d_c.h, d_v.h, d_x.h, d_y.h, 0, A.rows, A.nnz, None) == 1  # secsize 0
