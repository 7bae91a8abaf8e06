"""Pins the CPU oracle (oracle/mlem_oracle.c) and the host preprocessing that
shares its definitions, without a GPU.

* Transpose: bit-exact against the reference's own sptrans (golden vectors
  written by tests/golden/make_golden.py from oracle/_ref), against numpy's
  stable argsort and against scipy, at several thread counts.
* Partition, merge-path coordinates: brute force from their definitions.
* SpMV, column sums, initial image: scipy / sequential float32 / closed forms.
* MLEM: invariants that hold in exact arithmetic (diagonal system, the
  sensitivity-weighted sum, log-likelihood ascent, x0-scale invariance,
  zero rows and columns, nonnegativity) and a hand-worked 2x2 case.
"""
import os

import numpy as np
import pytest

import mlem_amd as M
import oracle as O
from matrices import KINDS, from_lengths, make

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_sptrans():
    z = np.load(os.path.join(GOLDEN, "sptrans_golden.npz"))
    out = []
    for name in z["names"]:
        rows, cols = (int(v) for v in z[f"{name}/shape"])
        out.append((str(name), rows, cols, z[f"{name}/rowptr"], z[f"{name}/col"], z[f"{name}/val"],
                    z[f"{name}/colptrT"], z[f"{name}/rowT"], z[f"{name}/valT"]))
    return out


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


# ------------------------------------------------------------------ transpose
@pytest.mark.parametrize("procs", [1, 2, 3, 8, 64])
def test_oracle_sptrans_matches_reference_golden(procs):
    for name, rows, cols, rowptr, col, val, colptrT, rowT, valT in golden_sptrans():
        cp, rt, vt = O.sptrans(rows, cols, rowptr, col, val, procs=procs)
        assert same_bits(cp, colptrT) and same_bits(rt, rowT) and same_bits(vt, valT), (name, procs)


@pytest.mark.parametrize("threads", [1, 2, 3, 8, 64])
def test_host_transpose_matches_reference_golden(threads):
    for name, rows, cols, rowptr, col, val, colptrT, rowT, valT in golden_sptrans():
        T = M.transpose(M.Csr(rows, cols, rowptr, col, val), threads=threads)
        assert same_bits(T.rowptr, colptrT) and same_bits(T.col, rowT) and same_bits(T.val, valT), (name, threads)


def stable_counting_sort_transpose(rows, cols, rowptr, col, val):
    """For each column, its entries in original nnz order: numpy's stable argsort on the column index."""
    order = np.argsort(col, kind="stable")
    row_of = np.repeat(np.arange(rows, dtype=np.uint32), np.diff(rowptr.astype(np.int64)))
    colptr = np.zeros(cols + 1, np.uint64)
    colptr[1:] = np.cumsum(np.bincount(col.astype(np.int64), minlength=cols))
    return colptr, row_of[order], val[order]


def test_golden_is_the_stable_counting_sort():
    # The definition pins the golden vectors too: a transposed operand or a
    # dropped stability requirement would fail here.
    for name, rows, cols, rowptr, col, val, colptrT, rowT, valT in golden_sptrans():
        cp, rt, vt = stable_counting_sort_transpose(rows, cols, rowptr, col, val)
        assert same_bits(cp, colptrT) and same_bits(rt, rowT) and same_bits(vt, valT), name


def test_transpose_matches_scipy_on_duplicate_free_matrices():
    for name, rows, cols, rowptr, col, val, colptrT, rowT, valT in golden_sptrans():
        if name == "dups" or name == "single_col":
            continue
        A = M.Csr(rows, cols, rowptr, col, val).to_scipy()
        T = A.T.tocsr()
        assert np.array_equal(T.indptr.astype(np.uint64), colptrT), name
        assert np.array_equal(T.indices.astype(np.uint32), rowT), name
        assert same_bits(T.data.astype(np.float32), valT), name


def test_transpose_random_matrices_all_paths_agree():
    rng = np.random.default_rng(7)
    mats = [make(k, seed=11) for k in KINDS]
    mats.append(from_lengths(rng.integers(0, 200, 700), 5000, seed=12))
    mats.append(from_lengths(rng.integers(0, 3, 10000), 3, seed=13, dup=True))
    for A in mats:
        ref = stable_counting_sort_transpose(A.rows, A.cols, A.rowptr, A.col, A.val)
        for t in (1, 3, 8):
            cp, rt, vt = O.sptrans(A.rows, A.cols, A.rowptr, A.col, A.val, procs=t)
            T = M.transpose(A, threads=t)
            for got in ((cp, rt, vt), (T.rowptr, T.col, T.val)):
                assert all(same_bits(a, b) for a, b in zip(got, ref))
        if O.ref_available():  # the reference's own code, where it was built here
            for t in (1, 4):
                got = O.ref_sptrans(A.rows, A.cols, A.rowptr, A.col, A.val, threads=t)
                assert all(same_bits(a, b) for a, b in zip(got, ref))


def test_transpose_without_values_and_twice_is_identity():
    A = make("ragged", seed=3)
    T = M.transpose(A, values=False)
    ref = stable_counting_sort_transpose(A.rows, A.cols, A.rowptr, A.col, A.val)
    assert same_bits(T.rowptr, ref[0]) and same_bits(T.col, ref[1])
    TT = M.transpose(M.transpose(A))  # rows sorted by column: A^T^T == A
    assert same_bits(TT.rowptr, A.rowptr) and same_bits(TT.col, A.col) and same_bits(TT.val, A.val)


# ------------------------------------------------------------------ partition
def partition_bruteforce(rowptr, rows, D):
    """segments[s] = first i in 0..rows with (double)rowptr[i] >= s*(double)nnz/D; segments[D] = rows."""
    nnz = float(rowptr[-1])
    seg = [0] * (D + 1)
    for s in range(D):
        seg[s] = next(i for i in range(rows + 1) if float(rowptr[i]) >= s * (nnz / D))
    seg[D] = rows
    return seg


def test_partition_against_definition():
    rng = np.random.default_rng(5)
    for trial in range(200):
        rows = int(rng.integers(1, 40))
        lens = rng.integers(0, 9, rows) * (rng.random(rows) < 0.7)
        rowptr = np.zeros(rows + 1, np.uint64)
        rowptr[1:] = np.cumsum(lens)
        nnz = int(rowptr[-1])
        for D in range(1, 9):
            seg, srows, snnz, off = O.partition(rowptr, nnz, rows, D)
            hseg, hrows, hnnz, hnnz64, hoff = M.partition(rowptr, nnz, rows, D)
            exp = partition_bruteforce(rowptr, rows, D)
            assert list(seg) == exp and list(hseg) == exp
            assert np.all(np.diff(seg.astype(np.int64)) >= 0)
            assert list(srows) == [exp[s + 1] - exp[s] for s in range(D)] == list(hrows)
            assert list(snnz) == [int(rowptr[exp[s + 1]] - rowptr[exp[s]]) for s in range(D)] == list(hnnz)
            assert list(hnnz64) == list(hnnz) and list(off) == [int(rowptr[exp[s]]) for s in range(D)] == list(hoff)
    seg, *_ = O.partition(np.array([0, 3, 5], np.uint64), 5, 2, 1)
    assert list(seg) == [0, 2]  # D = 1: the whole matrix


def test_partition_counts_truncate_to_32_bits_like_the_reference():
    rowptr = np.array([0, 2**32 + 5, 2**32 + 9], np.uint64)
    seg, srows, snnz, snnz64, off = M.partition(rowptr, int(rowptr[-1]), 2, 1)
    assert int(snnz64[0]) == 2**32 + 9 and int(snnz[0]) == 9


# ------------------------------------------------------------ merge path
def merge_bruteforce(rowptr, rows, d):
    """Walk d items of the merge of A = rowptr[1..rows] (row ends) and B = 0..nnz-1:
    take B[j] while j < rowptr[i+1], else A[i]; returns (i, j) after d items."""
    nnz = int(rowptr[-1])
    i = j = 0
    for _ in range(d):
        if i < rows and (j >= nnz or int(rowptr[i + 1]) <= j):
            i += 1
        else:
            j += 1
    return i, j


def test_merge_coordinates_against_bruteforce_merge():
    rng = np.random.default_rng(9)
    for trial in range(150):
        rows = int(rng.integers(1, 25))
        lens = rng.integers(0, 6, rows) * (rng.random(rows) < 0.8)
        rowptr = np.zeros(rows + 1, np.uint64)
        rowptr[1:] = np.cumsum(lens)
        total = rows + int(rowptr[-1])
        # Every thread starts below rows + nnz (kernel.cu:199-200); at d = rows + nnz
        # the reference's search is not entered for rows <= 1 (nexti = rows/2 = i = 0),
        # and no walk starts there.
        for d in range(total):
            assert O.merge_coord(rowptr, rows, d) == merge_bruteforce(rowptr, rows, d), (trial, d)
        if rows >= 2:
            assert O.merge_coord(rowptr, rows, total) == (rows, int(rowptr[-1]))


# ------------------------------------------------------------------ SpMV
def test_spmv_and_merge_spmv_against_scipy():
    rng = np.random.default_rng(4)
    for k in KINDS:
        A = make(k, seed=21)
        x = rng.random(A.cols).astype(np.float32)
        ref = A.to_scipy().astype(np.float64) @ x.astype(np.float64)
        y = O.spmv_f64(A.rowptr, A.col, A.val, x.astype(np.float64), A.rows)
        np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)
        for s in (1, 2, 3, 7, 32):
            ym = O.merge_spmv_f64(A.rowptr, A.col, A.val, x, A.rows, s)
            np.testing.assert_allclose(ym, ref, rtol=1e-12, atol=1e-12)


# ------------------------------------------------- sensitivity and init
def test_colsum_f64_against_scipy_and_f32_sequential_order():
    for k in ("ragged", "dups", "empty_rows", "tiny"):
        A = make(k, seed=2)
        s64 = O.colsum_f64(A.rowptr, A.col, A.val, A.rows, A.cols)
        np.testing.assert_allclose(s64, np.asarray(A.to_scipy().astype(np.float64).sum(axis=0)).ravel(),
                                   rtol=1e-12, atol=0)
        # float32, row order (one OpenMP thread of calcColumnSums)
        seq = np.zeros(A.cols, np.float32)
        for r in range(A.rows):
            for kk in range(int(A.rowptr[r]), int(A.rowptr[r + 1])):
                seq[A.col[kk]] = np.float32(seq[A.col[kk]] + A.val[kk])
        assert same_bits(O.colsum_f32(A.rowptr, A.col, A.val, A.rows, A.cols), seq)
        assert same_bits(M.colsum_f32(A), seq)


def test_initial_image_closed_form():
    rng = np.random.default_rng(1)
    for trial in range(20):
        rows, cols = int(rng.integers(1, 300)), int(rng.integers(1, 200))
        g = rng.integers(0, 1000, rows + 5).astype(np.int32)
        norm = (rng.random(cols) * 3).astype(np.float32)
        sum_norm = np.float32(0)
        for v in norm:
            sum_norm = np.float32(sum_norm + v)
        sum_g = int(np.sum(g[:rows].astype(np.int64)))
        sum_g = (sum_g + 2**31) % 2**32 - 2**31  # int accumulator
        init = np.float32(np.float32(sum_g) / sum_norm)
        assert np.float32(O.initial_image(g, rows, norm)) == init
        f0, hinit = M.initial_image(g, rows, norm)
        assert np.float32(hinit) == init and np.all(f0 == init)
    # the int sum wraps like the reference's
    g = np.array([2**31 - 1, 5], np.int32)
    f0, init = M.initial_image(g, 2, np.ones(1, np.float32))
    assert init == np.float32(-(2**31) + 4)


# ------------------------------------------------------------------ MLEM
def mlem(A, g, sens, f0, iters):
    return O.mlem_f64(A.rowptr, A.col, A.val, A.rows, A.cols, g, sens, f0, iters)


def test_mlem_hand_worked_2x2():
    # A = [[1, 2], [0, 3]], g = [6, 9], x0 = [1, 1], sens = [1, 5]:
    # fp = [3, 3], ratio = [2, 3], bp = [2, 4 + 9] = [2, 13], x1 = [2/1, 13/5].
    A = M.Csr(2, 2, np.array([0, 2, 3], np.uint64), np.array([0, 1, 1], np.uint32),
              np.array([1, 2, 3], np.float32))
    x1 = mlem(A, np.array([6, 9], np.int32), np.array([1.0, 5.0]), np.ones(2), 1)
    np.testing.assert_allclose(x1, [2.0, 2.6], rtol=1e-15)


def test_mlem_diagonal_converges_in_one_step():
    rng = np.random.default_rng(3)
    n = 50
    d = (rng.random(n) + 0.1).astype(np.float32)
    A = M.Csr(n, n, np.arange(n + 1, dtype=np.uint64), np.arange(n, dtype=np.uint32), d)
    g = rng.integers(1, 100, n).astype(np.int32)
    x1 = mlem(A, g, d.astype(np.float64), rng.random(n) + 0.5, 1)
    np.testing.assert_allclose(x1, g / d.astype(np.float64), rtol=1e-13)


def test_mlem_invariants_on_scanner_matrix(c1):
    A, g = c1["A"], c1["g"]
    sens = O.colsum_f64(A.rowptr, A.col, A.val, A.rows, A.cols)
    x = np.ones(A.cols)
    lls = [O.loglik_f64(A.rowptr, A.col, A.val, A.rows, g, x)]
    for k in range(8):
        fp = O.spmv_f64(A.rowptr, A.col, A.val, x, A.rows)
        x_next = mlem(A, g, sens, x, 1)
        # sum_j sens_j x^{k+1}_j = sum over rows with (A x^k)_i > 0 of g_i
        lhs = float(np.dot(sens, x_next))
        rhs = float(np.sum(g[fp > 0].astype(np.float64)))
        assert abs(lhs - rhs) <= 1e-10 * rhs
        assert np.all(x_next >= 0)
        x = x_next
        lls.append(O.loglik_f64(A.rowptr, A.col, A.val, A.rows, g, x))
    assert all(b >= a - 1e-9 * abs(a) for a, b in zip(lls, lls[1:]))
    assert lls[-1] > lls[0]
    # A uniform x0 scale does not change x^k for k >= 1.
    a = mlem(A, g, sens, np.full(A.cols, 0.37), 3)
    b = mlem(A, g, sens, np.full(A.cols, 91.0), 3)
    np.testing.assert_allclose(a, b, rtol=1e-12)


def test_mlem_zero_rows_and_columns():
    # An empty row with counts contributes nothing (fp = 0 -> ratio 0): the
    # image equals that of the matrix without the row.  A zero column has
    # sens = 0 and bp = 0, so its voxel becomes f * bp = 0 (no division).
    A = make("ragged", seed=8)
    rng = np.random.default_rng(8)
    g = rng.integers(1, 50, A.rows).astype(np.int32)
    lens = np.diff(A.rowptr.astype(np.int64))
    empty = np.flatnonzero(lens == 0)
    assert len(empty) > 0
    keep = np.flatnonzero(lens > 0)
    S = A.to_scipy()
    B = M.Csr.from_scipy(S[keep])
    sens = O.colsum_f64(A.rowptr, A.col, A.val, A.rows, A.cols)
    x0 = rng.random(A.cols) + 0.5
    xa = mlem(A, g, sens, x0, 4)
    xb = mlem(B, g[keep], sens, x0, 4)
    np.testing.assert_allclose(xa, xb, rtol=1e-13)
    zero_cols = np.flatnonzero(sens == 0)
    if len(zero_cols):
        assert np.all(xa[zero_cols] == 0)
    # add an explicit zero column and check it too
    C = M.Csr(A.rows, A.cols + 1, A.rowptr, A.col, A.val)
    sc = np.append(sens, 0.0)
    xc = mlem(C, g, sc, np.append(x0, 2.0), 1)
    assert xc[-1] == 0.0 and np.all(np.isfinite(xc))


def test_mlem_f32_cpu_baseline_matches_fp64(c1):
    A, AT, g = c1["A"], c1["AT"], c1["g"]
    s32 = M.colsum_f32(A)
    f0 = np.ones(A.cols, np.float32)
    ref = mlem(A, g, s32.astype(np.float64), f0, 20)
    for t in (1, 4):
        img, secs = O.mlem_f32_omp(A, AT, g, s32, f0, 20, t)
        assert np.linalg.norm(img - ref) / np.linalg.norm(ref) <= 1e-5
        assert secs > 0


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_mlem_f64_omp_is_bitwise_the_serial_oracle(c1, threads):
    """The OpenMP FP64 oracle (FP over rows of A, BP over rows of the stable transpose) forms every sum in the
    serial oracle's order, so each image is bitwise the serial one, for one image and for B of them."""
    A, AT, g = c1["A"], c1["AT"], c1["g"]
    sens = O.colsum_f64(A.rowptr, A.col, A.val, A.rows, A.cols)
    f0 = np.ones(A.cols)
    ref = mlem(A, g, sens, f0, 6)
    one = O.mlem_f64_omp(A, AT, g, sens, f0, 6, threads)
    assert same_bits(one, ref)
    rng = np.random.default_rng(threads)
    G = np.stack([g, rng.integers(0, 30, A.rows).astype(np.int32), np.roll(g, 7)])
    F0 = np.stack([f0, rng.random(A.cols) + 0.5, np.full(A.cols, 3.0)])
    many = O.mlem_f64_omp(A, AT, G, sens, F0, 6, threads)
    for b in range(3):
        assert same_bits(many[b], mlem(A, G[b], sens, F0[b], 6)), b


@pytest.mark.parametrize("kind", ["empty_rows", "dups", "heavy", "all_empty", "one_huge_row"])
def test_mlem_f64_omp_edge_matrices(kind):
    A = make(kind, seed=21)
    AT = M.transpose(A)
    rng = np.random.default_rng(21)
    g = rng.integers(0, 40, A.rows).astype(np.int32)
    sens = O.colsum_f64(A.rowptr, A.col, A.val, A.rows, A.cols)
    f0 = rng.random(A.cols) + 0.25
    assert same_bits(O.mlem_f64_omp(A, AT, g, sens, f0, 3, 4), mlem(A, g, sens, f0, 3))


def test_planted_mistakes_are_caught():
    """Each check above must fail for a plausible slip: here the slips are
    planted in numpy restatements of one iteration (not in the oracle) and
    the same invariants reject them."""
    A = make("ragged", seed=14)
    rng = np.random.default_rng(14)
    g = rng.integers(1, 40, A.rows).astype(np.int32)
    S = A.to_scipy().astype(np.float64)
    sens = np.asarray(S.sum(axis=0)).ravel()
    x = rng.random(A.cols) + 0.5

    def step(x, transpose_slip=False, drop_ratio=False, sens_slip=False):
        fp = S @ x
        r = np.where(fp != 0, g / np.where(fp != 0, fp, 1), 0.0)
        if drop_ratio:
            r = g.astype(np.float64)
        bp = (S.T @ r) if not transpose_slip else (S.T @ np.roll(r, 1))
        den = np.where(sens == 0, 1.0, sens) if not sens_slip else np.where(sens == 0, 1.0, sens) ** 2
        return x * bp / den

    good = mlem(A, g, sens, x, 1)
    np.testing.assert_allclose(step(x), good, rtol=1e-12)
    fp = S @ x
    rhs = float(np.sum(g[fp > 0]))
    for slip in ({"transpose_slip": True}, {"drop_ratio": True}, {"sens_slip": True}):
        bad = step(x, **slip)
        assert abs(float(np.dot(sens, bad)) - rhs) > 1e-6 * rhs, slip


def test_loglik_hand_worked():
    # A = [[1, 2], [0, 0], [0, 3]], f = [1, 2], g = [4, 5, 6]:
    # y = [5, 0, 6]: 4 log 5 - 5 + (-0) + 6 log 6 - 6.
    A = M.Csr(3, 2, np.array([0, 2, 2, 3], np.uint64), np.array([0, 1, 1], np.uint32),
              np.array([1, 2, 3], np.float32))
    ll = O.loglik_f64(A.rowptr, A.col, A.val, 3, np.array([4, 5, 6], np.int32), np.array([1.0, 2.0]))
    assert abs(ll - (4 * np.log(5) - 5 + 6 * np.log(6) - 6)) < 1e-12


# ------------------------------------------------------------ 16-bit values
def test_quantize_values_known_codes():
    """The 16-bit restatement against known codes, round-to-nearest-even ties
    and the closed-form error bounds."""
    f = np.float32
    # bf16: 1.0, a tie to even below, a tie to even above
    assert list(O.bf16_code(np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8], f))) == [0x3F80, 0x3F80, 0x3F82]
    assert O.bf16_value(np.array([0x3F81]))[0] == f(1 + 2 ** -7)
    # uf16 (4 exponent bits, 12 mantissa bits, 1.0 in the top binade)
    assert list(O.uf16_code(np.array([1.0, 0.75, 2 ** -15, 0.0, 2.0], f))) == [0xF000, 0xE800, 0, 0, 0xF000]
    t_tie_even, t_tie_up = f(0.5 * (1 + 2 ** -13)), f(0.5 * (1 + 3 * 2 ** -13))
    assert list(O.uf16_code(np.array([t_tie_even, t_tie_up]))) == [0xE000, 0xE002]
    assert O.uf16_value(np.array([0xE001]))[0] == f(0.5 * (1 + 2 ** -12))
    rng = np.random.default_rng(5)
    t = np.concatenate([rng.random(20000), 2.0 ** rng.uniform(-15, 0, 20000)]).astype(f)
    t = t[t >= 2 ** -15]
    assert np.max(np.abs(O.uf16_value(O.uf16_code(t)) - t) / t) <= 2 ** -13
    assert np.max(np.abs(O.bf16_value(O.bf16_code(t)) - t) / t) <= 2 ** -8
    # whole rows: exact where the ratios are binary fractions; scale 1 for an empty row
    A = M.Csr(3, 4, np.array([0, 3, 3, 5], np.uint64), np.array([0, 1, 3, 0, 2], np.uint32),
              np.array([4.0, 1.0, 0.5, 3.0, 1.0], np.float32))
    for kind in (1, 2, 3):
        q = O.quantize_values(A.rowptr, A.val, A.rows, kind)
        assert list(q[:3]) == [4.0, 1.0, 0.5], kind
        rel = abs(q[4] - 1.0)
        assert 0 < rel <= {1: 2 ** -8, 2: 2 ** -11, 3: 2 ** -13}[kind], (kind, rel)
