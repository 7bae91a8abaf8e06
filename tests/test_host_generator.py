"""The synthetic-data generator (host library): the reference's matrices came
from external tools, so these pin the generator by geometry and statistics.

* 2D ring, one ray per LOR: each row's lengths sum to the chord of the segment
  between the crystal faces clipped to the image square (Liang-Barsky).
* Determinism: the same seed gives the same bytes for any thread count and
  for any split of the rows; rows are sorted by column without duplicates.
* Poisson sampler: mean and variance over many streams.
* Sinogram: the expected counts are scale * A x with the requested total.
"""
import math

import numpy as np
import pytest

import mlem_amd as M


def small_3d():
    return M.geometry(n_rings=2, n_per_ring=64, radius=35.5, crystal=(1.0, 10.0, 10.0), ring_pitch=10.0,
                      grid=(32, 32, 8), voxel=(2.0, 2.0, 2.5), sub_rays=4)


def clip_length(p, q, half):
    """Length of segment pq inside the square [-half, half]^2 (Liang-Barsky)."""
    dx, dy = q[0] - p[0], q[1] - p[1]
    t0, t1 = 0.0, 1.0
    for pk, qk in ((-dx, p[0] + half), (dx, half - p[0]), (-dy, p[1] + half), (dy, half - p[1])):
        if pk == 0.0:
            if qk < 0.0:
                return 0.0
            continue
        t = qk / pk
        if pk < 0.0:
            t0 = max(t0, t)
        else:
            t1 = min(t1, t)
    return max(0.0, t1 - t0) * math.hypot(dx, dy)


def test_2d_ring_row_sums_are_clipped_chords():
    geo = M.RING2D_288()
    A = M.gen_system_matrix(geo, seed=1)
    n = 288
    assert A.rows == n * (n - 1) // 2 and A.cols == 64 * 64
    cs = np.concatenate([[0.0], np.cumsum(A.val.astype(np.float64))])
    rp = A.rowptr.astype(np.int64)
    sums = cs[rp[1:]] - cs[rp[:-1]]
    r = 0
    worst = 0.0
    for a in range(n):
        pa = (40.0 * math.cos(2 * math.pi * a / n), 40.0 * math.sin(2 * math.pi * a / n))
        for b in range(a + 1, n):
            pb = (40.0 * math.cos(2 * math.pi * b / n), 40.0 * math.sin(2 * math.pi * b / n))
            exp = clip_length(pa, pb, 32.0)
            got = float(sums[r])
            if exp > 1e-3:
                worst = max(worst, abs(got - exp) / exp)
            else:
                assert got < 1e-3
            r += 1
    assert worst < 2e-5


def test_rows_sorted_unique_and_positive():
    for geo in (M.RING2D_288(), small_3d()):
        A = M.gen_system_matrix(geo, seed=3)
        assert np.all(A.val > 0) and np.all(A.col < A.cols)
        rp = A.rowptr.astype(np.int64)
        d = np.diff(A.col.astype(np.int64))
        starts = np.zeros(A.nnz, bool)
        starts[rp[1:-1][rp[1:-1] < A.nnz]] = True
        assert np.all((d > 0) | starts[1:]), "columns not strictly increasing within a row"


def test_generation_is_deterministic_across_threads_and_row_splits():
    geo = small_3d()
    a = M.gen_system_matrix(geo, seed=7, threads=1)
    b = M.gen_system_matrix(geo, seed=7, threads=5)
    for x, y in ((a.rowptr, b.rowptr), (a.col, b.col), (a.val.view(np.uint32), b.val.view(np.uint32))):
        assert np.array_equal(x, y)
    lo, hi = 1000, 3100
    s = M.gen_system_matrix(geo, seed=7, row_begin=lo, row_end=hi, threads=3)
    assert s.rows == hi - lo
    o = int(a.rowptr[lo])
    assert np.array_equal(s.rowptr.astype(np.int64), a.rowptr[lo:hi + 1].astype(np.int64) - o)
    assert np.array_equal(s.col, a.col[o:int(a.rowptr[hi])])
    assert np.array_equal(s.val.view(np.uint32), a.val[o:int(a.rowptr[hi])].view(np.uint32))
    c = M.gen_system_matrix(geo, seed=8, threads=2)
    assert not (c.nnz == a.nnz and np.array_equal(c.val.view(np.uint32), a.val.view(np.uint32)))


@pytest.mark.parametrize("lam", [0.3, 4.0, 9.5, 10.5, 60.0, 2500.0])
def test_poisson_sampler_moments(lam):
    n = 40000
    x = np.array([M.poisson_sample(lam, 11, i) for i in range(n)], np.float64)
    assert x.min() >= 0
    se = math.sqrt(lam / n)
    assert abs(x.mean() - lam) < 5 * se
    # variance of the sample variance of a Poisson: (lam + 2 lam^2 (n/(n-1))) / n
    sv = math.sqrt((lam + 2 * lam * lam) / n)
    assert abs(x.var(ddof=1) - lam) < 6 * sv
    assert M.poisson_sample(0.0, 11, 5) == 0
    assert M.poisson_sample(lam, 11, 123) == M.poisson_sample(lam, 11, 123)


def test_phantom_and_sinogram():
    geo = small_3d()
    A = M.gen_system_matrix(geo, seed=2)
    x = M.gen_phantom(geo, seed=2, n_shapes=4, a_min=0.5, a_max=4.0)
    assert x.shape == (A.cols,) and np.all(x >= 0) and np.count_nonzero(x) > 0
    total = 2e6
    y, lam = M.gen_sinogram(A, x, total, seed=5, return_expected=True)
    ax = A.to_scipy().astype(np.float64) @ x.astype(np.float64)
    assert abs(lam.sum() - total) <= 1e-9 * total
    np.testing.assert_allclose(lam, ax * (total / ax.sum()), rtol=1e-6, atol=1e-9)
    assert abs(int(y.astype(np.int64).sum()) - total) < 6 * math.sqrt(total)
    y2 = M.gen_sinogram(A, x, total, seed=5, threads=3)
    assert np.array_equal(y, y2)
