"""Parity at full size on the 3D scanner matrix (two rings of 576 crystals,
662,976 LOR rows x 784,000 voxels of 140x140x40, 22 sub-rays per LOR,
1,527,217,536 nonzeros: 12.2 GB as f32 + u32), the size the headline
benchmark runs at, with the reference's 100 iterations:

* the device transpose, whole and in column ranges, bitwise the host
  transpose (itself bitwise the reference's sptrans, tests/test_oracle.py),
  and the FP64 sensitivity against the oracle's column sums;
* every FP32 reconstruction path (the blocked default, eager and replayed as a
  graph; the reference's merge-path FP with CSR-vector BP on A^T; CSR-vector
  both; the plain merge path both; the scatter BP) after 100 iterations
  within 1e-5 relative L2 of the FP64 oracle (OpenMP restatement, bitwise the
  serial one), and the FP64 log-likelihood against the oracle's and rising;
* batches of 8 sinograms (counts log-uniform 1e6..1e9) with UF16, FP16, BF16
  and f32 values at 4 and 8 per pass: each sinogram's relative L2 against
  the FP64 oracle after 100 iterations, printed; f32 values within 1e-5,
  UF16 and FP16 within 1e-3 (the bar; measured up to 5.1e-4 and 6.0e-4);
  BF16 does not meet 1e-3 on this matrix (half an ulp of 2^-8 per value,
  grown over 100 iterations: up to 8.6e-3) and is held to a regression
  bound;
* the serving loop over three resident matrices (the 2D ring, the mid-size
  3D matrix and this one) with UF16 values and graph-captured iterations:
  every request bitwise equal to the same request served alone, and within
  1e-3 of the FP64 oracle.

The oracle runs on the host's cores: on the GPU box (16 threads) this
module takes 3.5 minutes, most of it the FP64 oracle (47 s for one image,
82 s for 8 images at 100 iterations).
"""
import gc
import time

import numpy as np
import pytest

import mlem_amd as M
import oracle as O

pytestmark = pytest.mark.gpu

ITERS = 100


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def same_bits(a, b):
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def say(msg):
    print(f"[xl-c2 {time.strftime('%H:%M:%S')}] {msg}", flush=True)


def oracle(A, AT, G, sens, F0, iters, chunk=10):
    """O.mlem_f64_omp in chunks of `chunk` iterations (the state between iterations is the image alone, so the
    result is the same bits), printing progress."""
    F = F0
    t = time.time()
    for done in range(0, iters, chunk):
        F = O.mlem_f64_omp(A, AT, G, sens, F, min(chunk, iters - done))
        say(f"  FP64 oracle: {min(done + chunk, iters)}/{iters} iterations, {time.time() - t:.1f} s")
    return F


@pytest.fixture(scope="module")
def c2():
    t = time.time()
    geo = M.RING3D_1152()
    A = M.gen_system_matrix(geo, seed=1)
    x_true = M.gen_phantom(geo, seed=1, n_shapes=8, a_min=0.5, a_max=4.0)
    g = M.gen_sinogram(A, x_true, 1e8, seed=1)
    say(f"matrix {A.rows} x {A.cols}, nnz {A.nnz} in {time.time() - t:.1f} s")
    assert A.nnz == 1527217536
    m = M.ResidentMatrix(A)
    m.build_transpose()
    AT = m.download_transpose()
    s32, s64 = m.sensitivity()
    d = dict(geo=geo, A=A, AT=AT, g=g, m=m, s32=s32, s64=s64)
    say(f"resident, device transpose downloaded, {time.time() - t:.1f} s")
    yield d
    m.close()
    d.clear()
    gc.collect()


@pytest.fixture(scope="module")
def ref100(c2):
    A, AT, g = c2["A"], c2["AT"], c2["g"]
    f0, init = M.initial_image(g, A.rows, c2["s32"])  # the reference's x0 = sum g / sum sens
    t = time.time()
    ref = oracle(A, AT, g, c2["s64"], f0, ITERS)
    say(f"FP64 oracle, {ITERS} iterations: {time.time() - t:.1f} s")
    return dict(f0=f0, ref=ref)


def test_c2_device_transpose_bitwise_the_host_transpose(c2):
    A, AT = c2["A"], c2["AT"]
    t = time.time()
    H = M.transpose(A)
    say(f"host transpose {time.time() - t:.1f} s")
    assert same_bits(H.rowptr, AT.rowptr) and same_bits(H.col, AT.col) and same_bits(H.val, AT.val)
    # The same sort in column ranges, within 3 GB of scratch (the path of matrices whose one sort does not fit).
    with M.ResidentMatrix(A) as m2:
        ranges = m2.build_transpose(scratch_bytes=3 << 30)
        R = m2.download_transpose()
    say(f"device transpose in {ranges} column ranges")
    assert ranges > 1
    assert same_bits(H.rowptr, R.rowptr) and same_bits(H.col, R.col) and same_bits(H.val, R.val)


def test_c2_sensitivity_fp64_matches_oracle(c2):
    A = c2["A"]
    ref = O.colsum_f64(A.rowptr, A.col, A.val, A.rows, A.cols)
    np.testing.assert_allclose(c2["s64"], ref, rtol=1e-12, atol=0)
    assert np.array_equal(c2["s32"], ref.astype(np.float32))


PATHS = {
    # name: (opts keywords, extra attributes)
    "blocked": (dict(spmv=0, using_trans=0), {}),
    "blocked-graph": (dict(spmv=0, using_trans=0, use_graph=1), {}),
    "merge-fp+csr-vector-bp": (dict(spmv=0, using_trans=0, tile_width=1), {"bp_spmv": 2}),
    "csr-vector": (dict(spmv=1, using_trans=0), {}),
    "plain-merge": (dict(spmv=0, using_trans=0, tile_width=1, secsize_fp=5, secsize_bp=5), {}),
    "scatter-bp": (dict(spmv=1, using_trans=1), {}),
}


@pytest.mark.parametrize("path", list(PATHS))
def test_c2_fp32_100_iterations_vs_fp64_oracle(c2, ref100, path):
    kw, extra = PATHS[path]
    o = M.opts(**kw)
    for k, v in extra.items():
        setattr(o, k, v)
    t = time.perf_counter()
    img = M.mlem_run(c2["m"], c2["g"], ref100["f0"], ITERS, sens=c2["s32"], options=o)
    secs = time.perf_counter() - t
    err = rel_l2(img, ref100["ref"])
    say(f"{path}: {ITERS} iterations in {secs:.2f} s, rel-L2 vs FP64 oracle {err:.3e}")
    assert np.all(np.isfinite(img)) and np.all(img >= 0)
    assert err <= 1e-5


def test_c2_graph_replay_bitwise_eager(c2, ref100):
    a = M.mlem_run(c2["m"], c2["g"], ref100["f0"], 12, sens=c2["s32"], options=M.opts(spmv=0, using_trans=0))
    b = M.mlem_run(c2["m"], c2["g"], ref100["f0"], 12, sens=c2["s32"],
                   options=M.opts(spmv=0, using_trans=0, use_graph=1))
    assert same_bits(a, b)


def test_c2_loglik_fp64_and_ascent(c2, ref100):
    A, g = c2["A"], c2["g"]
    with M.Recon(c2["m"], g, ref100["f0"], sens=c2["s32"], options=M.opts(spmv=0, using_trans=0)) as r:
        lls = []
        for k in range(5):
            lls.append(r.loglik())
            if k in (0, 4):
                ref = O.loglik_f64(A.rowptr, A.col, A.val, A.rows, g, r.image().astype(np.float64))
                assert abs(lls[-1] - ref) <= 1e-6 * abs(ref)
            r.iterate(4)
    say("log-likelihood " + ", ".join(f"{v:.9e}" for v in lls))
    assert all(b >= a - 1e-9 * abs(a) for a, b in zip(lls, lls[1:]))
    assert lls[-1] > lls[0]


# ------------------------------------------------------------- batches (C4)
B = 8


@pytest.fixture(scope="module")
def c4(c2):
    A, AT, geo = c2["A"], c2["AT"], c2["geo"]
    rng = np.random.default_rng(2026)
    t = time.time()
    sinos = []
    for k in range(B):
        x = M.gen_phantom(geo, seed=1000 + k, n_shapes=4 + k % 5, a_min=0.5, a_max=4.0)
        sinos.append(M.gen_sinogram(A, x, float(10 ** rng.uniform(6, 9)), seed=2000 + k))
    G = np.stack(sinos)
    F0 = np.ones((B, A.cols), np.float32)
    ref = oracle(A, AT, G, c2["s64"], F0, ITERS)
    say(f"{B} sinograms and their FP64 oracle images ({ITERS} iterations) in {time.time() - t:.1f} s")
    return dict(G=G, F0=F0, ref=ref, counts=[int(s.sum()) for s in sinos])


VALUES = {"f32": 0, "bf16": 1, "fp16": 2, "uf16": 3}
# Bound on every sinogram's relative L2 against the FP64 oracle after 100 iterations: the 1e-3 bar where the type
# meets it, else a regression bound above what was measured (BF16 up to 8.6e-3).
BOUND = {"f32": 1e-5, "uf16": 1e-3, "fp16": 1e-3, "bf16": 2e-2}


@pytest.mark.parametrize("per_pass", [4, 8])
@pytest.mark.parametrize("values", ["uf16", "fp16", "bf16", "f32"])
def test_c4_batched_per_sinogram_error_vs_fp64(c2, c4, values, per_pass):
    with M.Batch(c2["m"], B, values=VALUES[values], per_pass=per_pass) as b:
        for k in range(B):
            b.set(k, c4["G"][k], c4["F0"][k])
        t = time.perf_counter()
        b.iterate(ITERS)
        b.sync()
        secs = time.perf_counter() - t
        errs = [rel_l2(b.image(k), c4["ref"][k]) for k in range(B)]
    say(f"C4 {values} {per_pass} per pass: {ITERS} iterations of {B} in {secs:.2f} s; rel-L2 vs FP64 per sinogram "
        + " ".join(f"{e:.2e}" for e in errs) + f" (counts {c4['counts']})")
    assert max(errs) <= BOUND[values]


# -------------------------------------------------------------- serving (C5)
@pytest.fixture(scope="module")
def mid():
    geo = M.RING3D_576_MID()
    A = M.gen_system_matrix(geo, seed=1)
    m = M.ResidentMatrix(A)
    m.build_transpose()
    AT = m.download_transpose()
    _, s64 = m.sensitivity()
    rng = np.random.default_rng(77)
    sinos = [M.gen_sinogram(A, M.gen_phantom(geo, seed=700 + j, n_shapes=3 + j % 4, a_min=0.5, a_max=4.0),
                            float(10 ** rng.uniform(5, 7.5)), seed=710 + j) for j in range(4)]
    d = dict(A=A, AT=AT, m=m, s64=s64, sinos=sinos, f0=np.ones(A.cols, np.float32))
    yield d
    m.close()
    d.clear()
    gc.collect()


def oracle_images(A, AT, sens, sinos, f0, iters):
    """FP64 oracle images of sinos[k] after iters[k] iterations: one OpenMP run over all of them, each image
    taken when its count is reached."""
    order = np.argsort(iters, kind="stable")
    G = np.stack(sinos)
    F = np.stack([np.asarray(f0, np.float64)] * len(sinos))
    out = [None] * len(sinos)
    done = 0
    for k in order:
        if iters[k] > done:
            F = oracle(A, AT, G, sens, F, iters[k] - done)
            done = iters[k]
        out[k] = F[k].copy()
    return out


def test_c5_serving_mix_uf16_graphs_alone_equals_mix(c1, c2, c4, mid):
    s1 = O.colsum_f64(c1["A"].rowptr, c1["A"].col, c1["A"].val, c1["A"].rows, c1["A"].cols)
    rng = np.random.default_rng(2029)
    c1_sinos = [c1["g"]] + [M.gen_sinogram(c1["A"], M.gen_phantom(c1["geo"], seed=900 + j, n_shapes=3 + j),
                                           float(10 ** rng.uniform(4, 6.5)), seed=910 + j) for j in range(3)]
    m1 = M.ResidentMatrix(c1["A"], c1["AT"])
    mats = [dict(m=m1, A=c1["A"], AT=c1["AT"], s64=s1, sinos=c1_sinos, f0=np.ones(c1["A"].cols, np.float32)),
            dict(m=mid["m"], A=mid["A"], AT=mid["AT"], s64=mid["s64"], sinos=mid["sinos"], f0=mid["f0"]),
            dict(m=c2["m"], A=c2["A"], AT=c2["AT"], s64=c2["s64"], sinos=list(c4["G"][:4]),
                 f0=np.ones(c2["A"].cols, np.float32))]
    kw = dict(values=VALUES["uf16"], per_pass=4, use_graph=1)
    # 14 requests, iterations uniform in 10..100, the second half submitted while the first is iterating
    plan = [(mi, j, int(rng.integers(10, 101))) for mi, n in ((0, 4), (1, 4), (2, 4)) for j in range(n)]
    plan += [(0, 1, int(rng.integers(10, 101))), (2, 3, int(rng.integers(10, 101)))]
    order = rng.permutation(len(plan))
    plan = [plan[k] for k in order]
    srv = None
    try:
        srv = M.Server(**kw)
        ids = [srv.add_matrix(d["m"], slots=4) for d in mats]
        rid, results, steps = {}, {}, 0
        for k in range(7):
            mi, j, it = plan[k]
            rid[k] = srv.submit(ids[mi], mats[mi]["sinos"][j], mats[mi]["f0"], it)
        while len(results) < len(plan):
            if steps == 20:
                for k in range(7, len(plan)):
                    mi, j, it = plan[k]
                    rid[k] = srv.submit(ids[mi], mats[mi]["sinos"][j], mats[mi]["f0"], it)
            srv.step()
            steps += 1
            for k in list(rid):
                r = srv.poll(rid[k], ids[plan[k][0]])
                if r is not None:
                    results[k] = r[0]
                    del rid[k]
            assert steps < 20000
        srv.close()
        srv = None
        say(f"C5: {len(plan)} requests over 3 matrices served in {steps} steps")
        # each request alone: a server holding only its matrix and this one request
        for k, (mi, j, it) in enumerate(plan):
            with M.Server(**kw) as one:
                i1 = one.add_matrix(mats[mi]["m"], slots=4)
                r1 = one.submit(i1, mats[mi]["sinos"][j], mats[mi]["f0"], it)
                got = None
                while got is None:
                    one.step()
                    got = one.poll(r1, i1)
            assert same_bits(results[k], got[0]), f"request {k} (matrix {mi}, {it} iterations) differs from alone"
        # FP64 bar, per matrix with one oracle run over its requests
        for mi, d in enumerate(mats):
            ks = [k for k in range(len(plan)) if plan[k][0] == mi]
            refs = oracle_images(d["A"], d["AT"], d["s64"], [d["sinos"][plan[k][1]] for k in ks], d["f0"],
                                 [plan[k][2] for k in ks])
            for k, ref in zip(ks, refs):
                err = rel_l2(results[k], ref)
                say(f"  request {k}: matrix {mi}, {plan[k][2]} iterations, rel-L2 vs FP64 {err:.2e}")
                assert err <= 1e-3
    finally:
        if srv is not None:
            srv.close()
        m1.close()
