"""16-bit matrix values and batched reconstructions (many sinograms over one
resident matrix) against the FP64 oracle, on the 2D ring matrix.

Bars: f32 values 1e-5 rel-L2; fp16 values 1e-3; bf16 values 2e-3 (8-bit
significand: the stored matrix differs from the FP32 one by up to 2^-9 per
value, each value rounded to nearest after the per-row scale; the products,
sums, ratio and update stay FP32); uf16 values 1e-4 (an unsigned 16-bit
float, 12 mantissa bits: every value within 2^-13 of itself); a slot's image
is bitwise the same alone and in a full batch."""
import itertools

import numpy as np
import pytest

import mlem_amd as M
import oracle as O

pytestmark = pytest.mark.gpu

ITERS = 30
BAR = {0: 1e-5, 1: 2e-3, 2: 1e-3, 3: 1e-4}


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def setup(c1):
    A = c1["A"]
    m = M.ResidentMatrix(A, c1["AT"])
    s32, s64 = m.sensitivity()
    rng = np.random.default_rng(12)
    # several sinograms: seeded phantoms, counts from 1e4 to 1e7
    sinos = []
    for k in range(10):
        x = M.gen_phantom(c1["geo"], seed=100 + k, n_shapes=3 + k % 4)
        sinos.append(M.gen_sinogram(A, x, float(10 ** rng.uniform(4, 7)), seed=200 + k))
    f0 = np.ones(A.cols, np.float32)
    refs = [O.mlem_f64(A.rowptr, A.col, A.val, A.rows, A.cols, g, s32.astype(np.float64), f0, ITERS) for g in sinos]
    yield dict(m=m, s32=s32, sinos=sinos, f0=f0, refs=refs)
    m.close()


@pytest.mark.parametrize("values", [1, 2, 3])
def test_single_reconstruction_with_16bit_values(c1, setup, values):
    bar = BAR[values]
    print(f"single values={values}", flush=True)
    o = M.opts(spmv=0, values=values)
    img = M.mlem_run(setup["m"], setup["sinos"][0], setup["f0"], ITERS, sens=setup["s32"], options=o)
    err = rel_l2(img, setup["refs"][0])
    print(f"  rel-L2 {err:.2e}", flush=True)
    assert err <= bar


@pytest.mark.parametrize("values", [1, 2, 3])
def test_16bit_values_are_the_quantized_matrix(c1, setup, values):
    """The 16-bit kernels compute MLEM on exactly the quantized matrix of the
    oracle's restatement (per-row scale, codes rounded to nearest even): one
    reconstruction and a batch agree with FP64 MLEM on that matrix (and its
    FP64 column sums as the sensitivity) to FP32 rounding (2e-6), far below
    their distance to the unquantized matrix's result."""
    A = c1["A"]
    q = O.quantize_values(A.rowptr, A.val, A.rows, values).astype(np.float32)
    sens = O.colsum_f64(A.rowptr, A.col, q, A.rows, A.cols).astype(np.float32).astype(np.float64)
    g, f0 = setup["sinos"][3], setup["f0"]
    ref_q = O.mlem_f64(A.rowptr, A.col, q, A.rows, A.cols, g, sens, f0, ITERS)
    one = M.mlem_run(setup["m"], g, f0, ITERS, options=M.opts(spmv=0, values=values))
    with M.Batch(setup["m"], 4, values=values, per_pass=4) as b:
        b.set(2, g, f0)
        b.iterate(ITERS)
        batched = b.image(2)
    far = rel_l2(one, setup["refs"][3])
    print(f"values={values}: vs quantized oracle {rel_l2(one, ref_q):.2e} / {rel_l2(batched, ref_q):.2e}, "
          f"vs unquantized {far:.2e}", flush=True)
    assert rel_l2(one, ref_q) <= 2e-6 and rel_l2(batched, ref_q) <= 2e-6


def test_16bit_values_need_the_blocked_path(setup):
    with pytest.raises(M.MlemError):
        M.Recon(setup["m"], setup["sinos"][0], setup["f0"], sens=setup["s32"], options=M.opts(spmv=1, values=1))


@pytest.mark.parametrize("per_pass,width,values,threads",
                         [(1, 1024, 0, 0), (4, 512, 0, 0), (4, 1024, 0, 0), (8, 512, 0, 0), (4, 512, 1, 0),
                          (4, 512, 2, 0), (8, 512, 2, 0), (4, 1024, 1, 0), (1, 1024, 2, 0), (4, 512, 0, 256),
                          (4, 512, 0, 512), (8, 512, 2, 256), (4, 1024, 1, 512), (4, 512, 3, 0), (8, 512, 3, 0),
                          (1, 1024, 3, 0)])
def test_batch_matches_oracle(setup, per_pass, width, values, threads):
    print(f"batch per_pass={per_pass} W={width} values={values} threads={threads}", flush=True)
    n = len(setup["sinos"])
    with M.Batch(setup["m"], n, values=values, per_pass=per_pass, tile_width=width, tile_threads=threads) as b:
        for k, g in enumerate(setup["sinos"]):
            b.set(k, g, setup["f0"])
        b.iterate(ITERS)
        b.sync()
        errs = [rel_l2(b.image(k), setup["refs"][k]) for k in range(n)]
    print("  worst rel-L2", max(errs), flush=True)
    assert max(errs) <= BAR[values]


@pytest.mark.parametrize("values", [0, 1])
def test_slot_alone_equals_slot_in_full_batch(setup, values):
    k = 6
    with M.Batch(setup["m"], 1, values=values, per_pass=4) as alone:
        alone.set(0, setup["sinos"][k], setup["f0"])
        alone.iterate(ITERS)
        a = alone.image(0)
    with M.Batch(setup["m"], len(setup["sinos"]), values=values, per_pass=4) as mix:
        for j, g in enumerate(setup["sinos"]):
            mix.set(j, g, setup["f0"])
        mix.iterate(ITERS)
        c = mix.image(k)
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32))


def test_deactivated_slot_is_frozen_and_slots_are_reusable(setup):
    with M.Batch(setup["m"], 5, per_pass=4) as b:
        for k in range(5):
            b.set(k, setup["sinos"][k], setup["f0"])
        b.iterate(5)
        frozen = b.image(2)
        b.deactivate(2)
        b.iterate(7)
        assert np.array_equal(b.image(2).view(np.uint32), frozen.view(np.uint32))
        # slot 4 is alone in the second pass: deactivating it skips that pass
        b.deactivate(4)
        last = b.image(4)
        b.iterate(3)
        assert np.array_equal(b.image(4).view(np.uint32), last.view(np.uint32))
        # reuse slot 2 for another sinogram: it restarts from f0
        b.set(2, setup["sinos"][7], setup["f0"])
        b.iterate(ITERS)
        assert rel_l2(b.image(2), setup["refs"][7]) <= 1e-5
        info = b.info()
        assert info["per_pass"] == 4 and info["passes"] == 2 and info["device_bytes"] > 0


def test_uf16_values_refuse_negative_matrices():
    A = M.Csr(3, 4, np.array([0, 2, 3, 5], np.uint64), np.array([0, 2, 1, 0, 3], np.uint32),
              np.array([1.0, 0.5, -0.25, 2.0, 1.0], np.float32))
    g = np.array([3, 1, 2], np.int32)
    f0 = np.ones(4, np.float32)
    with M.ResidentMatrix(A, M.transpose(A)) as m:
        with pytest.raises(M.MlemError):
            M.Batch(m, 4, values=3)
        with pytest.raises(M.MlemError):
            M.Recon(m, g, f0, options=M.opts(spmv=0, values=3))
        with M.Batch(m, 4, values=2) as b:  # fp16 keeps the sign
            b.set(0, g, f0)


def test_batch_argument_errors(setup, c1):
    with pytest.raises(M.MlemError):
        M.Batch(setup["m"], 4, per_pass=3)
    with pytest.raises(M.MlemError):
        M.Batch(setup["m"], 4, per_pass=8, tile_width=1024)
    with pytest.raises(M.MlemError):
        M.Batch(setup["m"], 4, per_pass=4, tile_threads=64)
    with M.ResidentMatrix(c1["A"]) as no_t:
        with pytest.raises(M.MlemError):
            M.Batch(no_t, 4)
    with M.Batch(setup["m"], 2) as b:
        with pytest.raises(M.MlemError):
            b.set(2, setup["sinos"][0], setup["f0"])


@pytest.mark.parametrize("per_pass,values", [(4, 0), (4, 3), (1, 0)])
def test_graph_replays_equal_eager_iterations(setup, per_pass, values):
    """use_graph: every mlem_batch_iterate call replays one captured
    iteration; after each call (eager image reads and host syncs between
    the replays) every slot is bitwise the eager batch's, also after a slot
    leaves, a pass empties (a new capture) and a slot is reused."""
    n = 6
    sinos, f0 = setup["sinos"], setup["f0"]
    with M.Batch(setup["m"], n, values=values, per_pass=per_pass) as eager, \
            M.Batch(setup["m"], n, values=values, per_pass=per_pass, use_graph=1) as graph:
        for b in (eager, graph):
            for k in range(n):
                b.set(k, sinos[k], f0)

        def step(iters, what):
            for b in (eager, graph):
                b.iterate(iters)
                b.sync()
            for k in range(n):
                a, c = eager.image(k), graph.image(k)
                assert np.array_equal(a.view(np.uint32), c.view(np.uint32)), (what, k)

        for r in range(4):
            step(1, f"replay {r}")
        step(3, "three replays in one call")
        for b in (eager, graph):
            b.deactivate(n - 1)
            b.deactivate(n - 2)  # with 4 per pass: the second pass empties
        step(2, "after deactivation")
        for b in (eager, graph):
            b.set(n - 1, sinos[7], f0)
        step(2, "after reuse")
        assert not np.array_equal(graph.image(0), f0)  # it did iterate


@pytest.mark.parametrize("per_pass,width,rows,groups", [(4, 512, 512, 0), (4, 1024, 2048, 1), (8, 512, 1024, 100000)])
def test_batch_edge_matrices(per_pass, width, rows, groups):
    """Adversarial matrices through the batched kernels (f32 values): empty
    rows and columns, one 60,000-entry row (carry chains across whole
    windows), heavy tails, duplicate columns, no entries, one row, rows with
    gaps; one run per row block up to one window per run; two passes, the
    second one partly filled; workgroups of 128, 256 and 512 work-items (4,
    2 and 1 groups of a window each).  Every slot within 1e-5 of the FP64
    oracle."""
    from matrices import make
    for kind, threads in itertools.product(["empty_rows", "one_huge_row", "heavy", "dups", "all_empty", "single_row",
                                            "row_gaps"], [128, 256, 512]):
        A = make(kind, seed=7)
        AT = M.transpose(A)
        rng = np.random.default_rng(8)
        slots = per_pass + 1
        sinos = [rng.integers(0, 30, A.rows).astype(np.int32) for _ in range(slots)]
        with M.ResidentMatrix(A, AT) as m:
            s32, _ = m.sensitivity()
            f0 = np.ones(A.cols, np.float32)
            print(f"batch edge {kind} per_pass={per_pass} W={width} R={rows} groups={groups} threads={threads}",
                  flush=True)
            with M.Batch(m, slots, values=0, per_pass=per_pass, tile_width=width, tile_rows=rows,
                         tile_groups=groups, tile_threads=threads) as b:
                for k in range(slots):
                    b.set(k, sinos[k], f0)
                b.iterate(8)
                b.sync()
                for k in range(slots):
                    ref = O.mlem_f64(A.rowptr, A.col, A.val, A.rows, A.cols, sinos[k], s32.astype(np.float64), f0, 8)
                    img = b.image(k)
                    assert np.linalg.norm(ref) == 0 and np.all(img == 0) or rel_l2(img, ref) <= 1e-5, (kind, k)


@pytest.mark.parametrize("per_pass,values", [(1, 0), (4, 2)])
def test_autotune_choice_reconstructs(setup, per_pass, values):
    """The autotuner returns a shape among its candidates; reconstructions with it meet the bars."""
    t = setup["m"].autotune(per_pass=per_pass, values=values)
    print(f"autotune per_pass={per_pass} values={values}: {t}", flush=True)
    assert t["candidates"] >= 2 and t["fp_ms"] > 0 and t["bp_ms"] > 0
    assert t["tile_width"] in (512, 1024, 2048) and t["tile_rows"] in (512, 1024, 2048, 4096)
    assert t["tile_threads"] == 0 if per_pass == 1 else t["tile_threads"] in (256, 512)
    if per_pass == 1:
        o = M.opts(spmv=0, tile_width=t["tile_width"], tile_rows=t["tile_rows"], tile_groups=t["tile_groups"])
        img = M.mlem_run(setup["m"], setup["sinos"][0], setup["f0"], ITERS, sens=setup["s32"], options=o)
        assert rel_l2(img, setup["refs"][0]) <= BAR[0]
    else:
        with M.Batch(setup["m"], per_pass, values=values, per_pass=per_pass, tile_width=t["tile_width"],
                     tile_rows=t["tile_rows"], tile_groups=t["tile_groups"], tile_threads=t["tile_threads"]) as b:
            for k in range(per_pass):
                b.set(k, setup["sinos"][k], setup["f0"])
            b.iterate(ITERS)
            b.sync()
            for k in range(per_pass):
                assert rel_l2(b.image(k), setup["refs"][k]) <= BAR[values]
