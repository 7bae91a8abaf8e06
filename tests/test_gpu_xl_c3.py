"""The 100 GB-class matrix at full size (662,976 LOR rows x 6,272,000 voxels
of 280x280x80, 128 sub-rays per LOR, 1.65e10 nonzeros: 132 GB per CSR
copy), resident as the tiles of A and A^T only, through the reference's 20
iterations, checked against what MLEM fixes in exact arithmetic:

* the sensitivity-weighted image sum: sum_j sens_j x_j^{k+1} equals the counts
  of the rows with (A x^k)_i > 0, which from a positive x^0 are the non-empty
  rows with counts (a row with counts keeps a positive forward projection);
  summed in FP64 on the host over the FP32 image and the sensitivity the
  update divides by, within 1e-5;
* the Poisson log-likelihood, in FP64 on the GPU, rising every iteration;
* nonnegative, finite images.

The FP64 oracle is not run here: A and A^T do not fit the host memory of one
process beside each other.  About 2 minutes on the GPU box, most of it
generating the matrix on the host (132 GB) and building the tiles.
"""
import gc
import time

import numpy as np
import pytest

import mlem_amd as M

pytestmark = pytest.mark.gpu

ITERS = 20


def say(msg):
    print(f"[xl-c3 {time.strftime('%H:%M:%S')}] {msg}", flush=True)


def test_c3_tiles_only_fp64_invariants():
    t = time.time()
    geo = M.RING3D_1152_FINE()
    A = M.gen_system_matrix(geo, seed=1)
    x_true = M.gen_phantom(geo, seed=1, n_shapes=8, a_min=0.5, a_max=4.0)
    g = M.gen_sinogram(A, x_true, 1e8, seed=1)
    say(f"matrix {A.rows} x {A.cols}, nnz {A.nnz} ({A.nnz * 8 / 1e9:.0f} GB) in {time.time() - t:.1f} s")
    nonempty = np.diff(A.rowptr.astype(np.int64)) > 0
    counted = float(np.sum(g[nonempty], dtype=np.float64))
    m = M.ResidentMatrix.tiles_only(A)
    say(f"tiles of A and A^T resident: {m.info()['device_bytes'] / 1e9:.1f} GB, {time.time() - t:.1f} s")
    del A
    gc.collect()
    try:
        s32, s64 = m.sensitivity()
        assert np.all(s32 >= 0) and np.all(np.isfinite(s64))
        f0 = np.ones(m.cols, np.float32)
        with M.Recon(m, g, f0, options=M.opts(spmv=0, using_trans=0)) as r:
            lls = [r.loglik()]
            for k in range(ITERS):
                r.iterate(1)
                img = r.image()
                assert np.all(np.isfinite(img)) and np.all(img >= 0), k
                lhs = float(np.dot(s32.astype(np.float64), img.astype(np.float64)))
                rel = abs(lhs - counted) / counted
                lls.append(r.loglik())
                say(f"iteration {k + 1}: sum sens*x = {lhs:.9e} (counts {counted:.9e}, {rel:.1e}), "
                    f"log-likelihood {lls[-1]:.10e}")
                assert rel <= 1e-5, k
        assert all(b >= a - 1e-9 * abs(a) for a, b in zip(lls, lls[1:])) and lls[-1] > lls[0]
    finally:
        m.close()
