"""The serving loop (continuous batching over resident matrices): requests of
different iteration counts on two matrices join and leave their batches
between iterations; each result is bitwise the image the same request gets
alone in a batch of its own, and within the FP64-oracle bar (1e-5 rel-L2 for
f32 values; 1e-3 for UF16 values, served with graph-captured iterations as
the serving benchmark runs them).  Also: the HBM budget refuses a matrix that does not fit and
admits requests by the memory of their slots' passes (waiting while a new
pass does not fit, freeing idle passes of other matrices for them), more
requests than slots queue, and bad arguments are reported."""
import numpy as np
import pytest

import mlem_amd as M
import oracle as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def ring(n_per_ring, radius, grid):
    return M.geometry(1, n_per_ring, radius, (0.0, 0.0, 0.0), 0.0, (grid, grid, 1), (1.0, 1.0, 1.0), 0)


@pytest.fixture(scope="module")
def served(c1):
    """Two resident matrices (288- and 192-crystal rings), seeded sinograms for each."""
    mats = []
    for k, geo in enumerate((c1["geo"], ring(192, 30.0, 48))):
        A = c1["A"] if k == 0 else M.gen_system_matrix(geo, seed=3)
        AT = c1["AT"] if k == 0 else M.transpose(A)
        m = M.ResidentMatrix(A, AT)
        s32, _ = m.sensitivity()
        rng = np.random.default_rng(40 + k)
        sinos = [M.gen_sinogram(A, M.gen_phantom(geo, seed=300 + 10 * k + j, n_shapes=3 + j % 3),
                                float(10 ** rng.uniform(4, 6.5)), seed=400 + 10 * k + j) for j in range(6)]
        mats.append(dict(A=A, m=m, s32=s32, sinos=sinos, f0=np.ones(A.cols, np.float32)))
    yield mats
    for d in mats:
        d["m"].close()


def alone(d, g, iters, per_pass, values=0, use_graph=0):
    with M.Batch(d["m"], 1, per_pass=per_pass, values=values, use_graph=use_graph) as b:
        b.set(0, g, d["f0"])
        b.iterate(iters)
        b.sync()
        return b.image(0)


# (per pass, values, graph, rel-L2 bar against the FP64 oracle of the f32 matrix): f32 values to 1e-5; UF16 values
# (16-bit codes, the quantized matrix) to the 1e-3 bar of 16-bit reconstructions, with graph-captured iterations as
# the serving benchmark runs them
SERVE_CASES = [(1, 0, 0, 1e-5), (4, 0, 0, 1e-5), (4, 3, 1, 1e-3), (8, 3, 1, 1e-3)]


@pytest.mark.parametrize("per_pass,values,use_graph,bar", SERVE_CASES)
def test_mixed_requests_match_alone_and_oracle(served, per_pass, values, use_graph, bar):
    print(f"serve per_pass={per_pass} values={values} graph={use_graph}", flush=True)
    srv = M.Server(per_pass=per_pass, values=values, use_graph=use_graph)
    ids = [srv.add_matrix(d["m"], slots=5) for d in served]
    # 12 requests, iteration counts 3..20, submitted in two waves so that the
    # second wave joins batches that are already iterating
    plan = [(k % 2, k // 2, [7, 20, 3, 12, 9, 15][k // 2] + (k % 2)) for k in range(12)]
    rid = {}
    for k, (mi, j, it) in enumerate(plan[:7]):
        rid[k] = srv.submit(ids[mi], served[mi]["sinos"][j], served[mi]["f0"], it)
    results, steps = {}, 0
    while len(results) < len(plan):
        if steps == 4:
            for k, (mi, j, it) in enumerate(plan[7:], start=7):
                rid[k] = srv.submit(ids[mi], served[mi]["sinos"][j], served[mi]["f0"], it)
        active, queued = srv.step()
        steps += 1
        assert active <= 10
        for k in list(rid):
            r = srv.poll(rid[k], ids[plan[k][0]])
            if r is not None:
                results[k] = r
                del rid[k]
        assert steps < 200
    for k, (mi, j, it) in enumerate(plan):
        d = served[mi]
        img, lat = results[k]
        assert lat > 0
        ref_alone = alone(d, d["sinos"][j], it, per_pass, values, use_graph)
        assert np.array_equal(img.view(np.uint32), ref_alone.view(np.uint32)), f"request {k} differs from alone"
        A = d["A"]
        ref = O.mlem_f64(A.rowptr, A.col, A.val, A.rows, A.cols, d["sinos"][j], d["s32"].astype(np.float64),
                         d["f0"], it)
        err = rel_l2(img, ref)
        print(f"  request {k}: matrix {mi} iters {it} rel-L2 {err:.2e} latency {lat * 1e3:.1f} ms", flush=True)
        assert err <= bar
    info = srv.info()
    assert info["matrices"] == 2 and info["steps"] == steps
    srv.close()


@pytest.mark.parametrize("values,use_graph", [(0, 0), (3, 1)])
def test_survivors_move_to_fewer_passes(served, values, use_graph):
    """Requests of 3 and 12 iterations alternate over 8 slots (two passes of 4): after the third iteration each pass
    holds two survivors; those of the second pass move into the first, so the other 9 iterations sweep one pass
    instead of two (15 pass sweeps instead of 24), and every image is bitwise the one its request gets alone."""
    d = served[0]
    srv = M.Server(per_pass=4, values=values, use_graph=use_graph)
    mid = srv.add_matrix(d["m"], slots=8)
    iters = [3, 12, 3, 12, 3, 12, 3, 12]
    rids = [srv.submit(mid, d["sinos"][k % 6], d["f0"], it) for k, it in enumerate(iters)]
    done = {}
    for _ in range(200):
        srv.step()
        for k, r in enumerate(rids):
            if k not in done:
                out = srv.poll(r, mid)
                if out is not None:
                    done[k] = out[0]
        if len(done) == len(rids):
            break
    assert len(done) == len(rids)
    st = srv.stats()
    print(f"values={values} graph={use_graph}: {st}", flush=True)
    assert st["slot_moves"] == 2 and st["pass_sweeps"] == 15
    for k, it in enumerate(iters):
        ref = alone(d, d["sinos"][k % 6], it, 4, values, use_graph)
        assert np.array_equal(done[k].view(np.uint32), ref.view(np.uint32)), f"request {k} differs from alone"
    srv.close()


def test_more_requests_than_slots_queue(served):
    d = served[1]
    srv = M.Server(per_pass=4)
    mid = srv.add_matrix(d["m"], slots=2)
    rids = [srv.submit(mid, d["sinos"][j], d["f0"], 4) for j in range(5)]
    seen = []
    done = {}
    while len(done) < 5:
        active, queued = srv.step()
        seen.append((active, queued))
        for r in rids:
            if r not in done:
                out = srv.poll(r, mid)
                if out is not None:
                    done[r] = out[0]
        assert len(seen) < 50
    assert seen[0] == (2, 3)
    assert max(a for a, _ in seen) <= 2
    for j, r in enumerate(rids):
        assert np.array_equal(done[r].view(np.uint32), alone(d, d["sinos"][j], 4, 4).view(np.uint32))
    srv.close()


def test_budget_refuses_a_matrix_that_does_not_fit(served):
    d0, d1 = served
    srv = M.Server(hbm_budget_bytes=1)
    with pytest.raises(M.MlemError) as e:
        srv.add_matrix(d0["m"], slots=4)
    assert "budget" in str(e.value)
    assert srv.info()["used"] == 0 and srv.info()["matrices"] == 0
    srv.close()
    # a budget that holds the small matrix, its batch and one pass but not both matrices
    need1, pass1 = footprint(d1, 4)
    srv = M.Server(hbm_budget_bytes=need1 + pass1 + 1024)
    srv.add_matrix(d1["m"], slots=4)
    assert srv.info()["used"] == need1
    with pytest.raises(M.MlemError):
        srv.add_matrix(d0["m"], slots=4)
    assert srv.info()["matrices"] == 1
    srv.close()


def footprint(d, per_pass):
    """(matrix + batch bytes a server accounts when the matrix is added, bytes of one pass of slots)."""
    with M.Server(per_pass=per_pass) as probe:
        mid = probe.add_matrix(d["m"], slots=per_pass)
        need = probe.info()["used"]
        rid = probe.submit(mid, d["sinos"][0], d["f0"], 1)
        while probe.poll(rid, mid) is None:
            probe.step()
        return need, probe.info()["used"] - need


def drive(srv, submitted, budget):
    """Steps srv until every (request id, matrix id) in `submitted` is done; checks the budget after every step.
    Returns ({request id: image}, most requests active at once)."""
    done, most = {}, 0
    for _ in range(5000):
        active, _ = srv.step()
        most = max(most, active)
        assert srv.info()["used"] <= budget
        for rid, mid in submitted:
            if rid not in done:
                out = srv.poll(rid, mid)
                if out is not None:
                    done[rid] = out[0]
        if len(done) == len(submitted):
            return done, most
    raise AssertionError("requests did not finish")


def test_budget_admits_requests_by_pass(served):
    """Slots beyond what the budget holds stay unused: with room for one pass of 4, at most 4 of 8 slots run,
    the other requests wait, and each result is the same as alone."""
    d = served[1]
    need, pas = footprint(d, 4)
    budget = need + pas + pas // 2
    srv = M.Server(hbm_budget_bytes=budget, per_pass=4)
    mid = srv.add_matrix(d["m"], slots=8)
    subs = [(srv.submit(mid, d["sinos"][j], d["f0"], 3 + j), mid) for j in range(6)]
    done, most = drive(srv, subs, budget)
    assert most == 4
    for j, (rid, _) in enumerate(subs):
        assert np.array_equal(done[rid].view(np.uint32), alone(d, d["sinos"][j], 3 + j, 4).view(np.uint32))
    srv.close()


def test_budget_frees_idle_passes_of_other_matrices(served):
    """Two matrices and room for one pass between them: the second matrix's requests wait until the first
    matrix's pass has no active slot, which is then freed for them."""
    (n0, p0), (n1, p1) = footprint(served[0], 4), footprint(served[1], 4)
    budget = n0 + n1 + max(p0, p1)
    srv = M.Server(hbm_budget_bytes=budget, per_pass=4)
    ids = [srv.add_matrix(d["m"], slots=4) for d in served]
    subs = [(srv.submit(ids[k], served[k]["sinos"][j], served[k]["f0"], 4), ids[k]) for k in (0, 1) for j in (0, 1)]
    done, _ = drive(srv, subs, budget)
    for (rid, mid), (k, j) in zip(subs, [(k, j) for k in (0, 1) for j in (0, 1)]):
        d = served[k]
        assert np.array_equal(done[rid].view(np.uint32), alone(d, d["sinos"][j], 4, 4).view(np.uint32))
    srv.close()


def test_bad_requests_are_reported(served):
    d = served[1]
    srv = M.Server()
    mid = srv.add_matrix(d["m"], slots=1)
    with pytest.raises(M.MlemError):
        srv.submit(mid + 1, d["sinos"][0], d["f0"], 3)
    with pytest.raises(M.MlemError):
        srv.submit(mid, d["sinos"][0], d["f0"], 0)
    with pytest.raises(M.MlemError):
        srv.poll(12345, mid)
    # an idle server steps without work
    assert srv.step() == (0, 0)
    srv.close()
