"""Batched layouts with a row order (mlem_matrix_set_row_order) on the 2D
ring matrix.

* The identity order builds the same tiles as no order: images bitwise equal.
* An order P on A gives bitwise the images of a batch on the matrix whose
  rows were permuted by P on the host (with the sinogram permuted alike):
  the device reorder, the transpose of the reordered matrix and the placing
  of each count are exact.
* Against the FP64 oracle the images meet the same bars as without an order;
  a slot alone equals it in a full batch; served requests likewise.
* Refusals: not a permutation, a matrix in use, a tiles-only matrix."""
import numpy as np
import pytest

import mlem_amd as M
import oracle as O

pytestmark = pytest.mark.gpu

ITERS = 30
BAR = {0: 1e-5, 3: 1e-4}


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def rows_in_order(A, order):
    """A with row i in place order[i] (host restatement of the device reorder)."""
    inv = np.argsort(order)
    lens = np.diff(A.rowptr.astype(np.int64))[inv]
    rowptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    idx = np.concatenate([np.arange(A.rowptr[i], A.rowptr[i + 1], dtype=np.int64) for i in inv])
    return M.Csr(A.rows, A.cols, rowptr, A.col[idx], A.val[idx])


@pytest.fixture(scope="module")
def ring(c1):
    A, AT = c1["A"], c1["AT"]
    order = M.lor_order(A, tuple(c1["geo"].grid))
    m = M.ResidentMatrix(A, AT)
    s32, _ = m.sensitivity()
    m.close()
    rng = np.random.default_rng(31)
    sinos = []
    for k in range(6):
        x = M.gen_phantom(c1["geo"], seed=300 + k, n_shapes=3 + k % 3)
        sinos.append(M.gen_sinogram(A, x, float(10 ** rng.uniform(4, 7)), seed=400 + k))
    f0 = np.ones(A.cols, np.float32)
    refs = [O.mlem_f64(A.rowptr, A.col, A.val, A.rows, A.cols, g, s32.astype(np.float64), f0, ITERS) for g in sinos]
    return dict(A=A, AT=AT, order=order, sinos=sinos, f0=f0, refs=refs)


def run_batch(m, sinos, f0, values, per_pass, iters=ITERS):
    with M.Batch(m, len(sinos), values=values, per_pass=per_pass) as b:
        for k, g in enumerate(sinos):
            b.set(k, g, f0)
        b.iterate(iters)
        b.sync()
        return [b.image(k) for k in range(len(sinos))]


@pytest.mark.parametrize("values,per_pass", [(0, 4), (3, 4), (3, 8)])
def test_identity_order_is_bitwise_no_order(ring, values, per_pass):
    A, AT = ring["A"], ring["AT"]
    print(f"identity order values={values} per_pass={per_pass}", flush=True)
    with M.ResidentMatrix(A, AT) as plain, M.ResidentMatrix(A, AT) as ident:
        ident.set_row_order(np.arange(A.rows, dtype=np.uint32))
        a = run_batch(plain, ring["sinos"], ring["f0"], values, per_pass, iters=10)
        b = run_batch(ident, ring["sinos"], ring["f0"], values, per_pass, iters=10)
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


@pytest.mark.parametrize("values,per_pass", [(0, 4), (3, 4), (0, 8)])
def test_order_is_bitwise_the_host_reordered_matrix(ring, values, per_pass):
    A, order = ring["A"], ring["order"]
    print(f"order vs host-reordered values={values} per_pass={per_pass}", flush=True)
    Ap = rows_in_order(A, order)
    sinos_p = []
    for g in ring["sinos"]:
        gp = np.empty_like(g)
        gp[order] = g
        sinos_p.append(gp)
    with M.ResidentMatrix(A, ring["AT"]) as m, M.ResidentMatrix(Ap, M.transpose(Ap)) as mp:
        m.set_row_order(order)
        a = run_batch(m, ring["sinos"], ring["f0"], values, per_pass, iters=10)
        b = run_batch(mp, sinos_p, ring["f0"], values, per_pass, iters=10)
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


@pytest.mark.parametrize("values,per_pass", [(0, 4), (0, 8), (3, 4), (3, 8)])
def test_ordered_batch_matches_oracle_and_alone(ring, values, per_pass):
    print(f"ordered batch values={values} per_pass={per_pass}", flush=True)
    with M.ResidentMatrix(ring["A"], ring["AT"]) as m:
        m.set_row_order(ring["order"])
        imgs = run_batch(m, ring["sinos"], ring["f0"], values, per_pass)
        errs = [rel_l2(x, r) for x, r in zip(imgs, ring["refs"])]
        print(f"  worst rel-L2 {max(errs):.2e}", flush=True)
        assert max(errs) <= BAR[values]
        alone = run_batch(m, [ring["sinos"][4]], ring["f0"], values, per_pass)[0]
        assert np.array_equal(alone.view(np.uint32), imgs[4].view(np.uint32))


def test_ordered_serving_equals_alone(ring):
    with M.ResidentMatrix(ring["A"], ring["AT"]) as m:
        m.set_row_order(ring["order"])
        alone = [run_batch(m, [g], ring["f0"], 3, 4, iters=12)[0] for g in ring["sinos"][:3]]
        with M.Server(values=3, per_pass=4) as srv:
            mid = srv.add_matrix(m, slots=4)
            rids = [srv.submit(mid, g, ring["f0"], 12) for g in ring["sinos"][:3]]
            done = {}
            for _ in range(2000):
                srv.step()
                for r in rids:
                    if r not in done:
                        out = srv.poll(r, mid)
                        if out is not None:
                            done[r] = out[0]
                if len(done) == len(rids):
                    break
        assert len(done) == len(rids)
        for r, a in zip(rids, alone):
            assert np.array_equal(done[r].view(np.uint32), a.view(np.uint32))


def test_refusals(ring):
    A = ring["A"]
    with M.ResidentMatrix(A, ring["AT"]) as m:
        dup = np.arange(A.rows, dtype=np.uint32)
        dup[5] = 6
        with pytest.raises(M.MlemError):
            m.set_row_order(dup)
        big = np.arange(A.rows, dtype=np.uint32)
        big[0] = A.rows
        with pytest.raises(M.MlemError):
            m.set_row_order(big)
        with pytest.raises(ValueError):
            m.set_row_order(np.arange(A.rows - 1, dtype=np.uint32))
        with M.Batch(m, 1, per_pass=4) as b:
            with pytest.raises(M.MlemError):
                m.set_row_order(ring["order"])
            b.set(0, ring["sinos"][0], ring["f0"])
            b.iterate(2)
        m.set_row_order(ring["order"])
        m.set_row_order(None)  # cleared: batches take A's rows as stored again
        plain = run_batch(m, [ring["sinos"][1]], ring["f0"], 0, 4, iters=5)[0]
    with M.ResidentMatrix(A, ring["AT"]) as m2:
        again = run_batch(m2, [ring["sinos"][1]], ring["f0"], 0, 4, iters=5)[0]
    assert np.array_equal(plain.view(np.uint32), again.view(np.uint32))
    t = M.ResidentMatrix.tiles_only(A, per_pass=4)
    try:
        with pytest.raises(M.MlemError):
            t.set_row_order(ring["order"])
    finally:
        t.close()
