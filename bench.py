#!/usr/bin/env python3
"""MLEM benchmark on one MI355X.

A step is one MLEM iteration (forward projection + ratio, back projection +
update) over the resident system matrix.  Default workload: the MADPET-II
scale 3D scanner (two rings of 576 crystals, 662,976 LOR rows x 784,000
voxels of 140x140x40, 22 jittered sub-rays per LOR: 1.53e9 nonzeros, 12.2 GB
as f32 + u32), a seeded ellipsoid phantom and Poisson counts totalling 1e8,
generated on the host from a seed (no data files exist for it).

Prints one JSON line: iterations/s over the timed steps, the dominant
kernel's achieved bandwidth against the 8 TB/s HBM peak (HIP events on the
reconstruction's stream), the FP32 OpenMP CPU loop timed on the host's
physical cores (3 runs of 1 warm-up + k iterations, median; k = the
workload's 100 iterations, more on small matrices so that the three runs
take about 25 s), and a witness: the GPU image after 100 iterations against
the FP64 OpenMP oracle.

Algorithmic bytes (f32 values, u32 indices, u64 row pointers):
  FP SpMV kernel  8*nnz + 8*(rows+1) + 4*cols      (A and x)
  BP SpMV kernel  8*nnz + 8*(cols+1) + 4*rows      (A^T and the ratio)
  iteration       16*nnz + 20*rows + 24*cols + 16  (both, with g, sens, f r/w)
`roofline.achieved` and `algorithmic_gbs` count these bytes.  The blocked
kernels read 6 bytes per stored entry (plus padding and the x segments), so
their algorithmic rate exceeds the HBM rate they draw: `roofline.traffic`
holds the measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE,
scaled by the same counters on a copy kernel) from profiles/traffic.json when
it was measured for this workload, and `roofline.drawn_gbs` / `drawn_frac`
the rate and fraction of peak those bytes make.
"""
import argparse
import json
import os
import sys
import time

# The CPU baseline runs one OpenMP thread per physical core, bound close (set
# before any library that links the OpenMP runtime is loaded).
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "bachelorarbeit-mlem_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "MLEM iters/s + HBM GB/s (12GB matrix); batched recon/s, p99 latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


_AFFINITY = sorted(os.sched_getaffinity(0))  # before the OpenMP runtime binds this thread to one place


def host_cpus():
    """(logical CPUs this process may run on, physical cores among them) from the affinity it started with and
    the sysfs topology."""
    cpus = _AFFINITY
    cores = set()
    for c in cpus:
        try:
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            with open(base + "physical_package_id") as f1, open(base + "core_id") as f2:
                cores.add((f1.read().strip(), f2.read().strip()))
        except OSError:
            cores.add(("cpu", str(c)))
    return len(cpus), len(cores)


def baseline_threads():
    """Threads of the CPU baseline: the physical cores this process may run on, at most OMP_NUM_THREADS (the
    host's CPU share) when that is set."""
    _, phys = host_cpus()
    cap = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(phys, int(cap))) if cap.isdigit() and int(cap) > 0 else phys


def cpu_baseline_f32(O, A, AT, g, s32, f0, threads, budget_s=25.0, runs=3, min_iters=100):
    """The FP32 OpenMP CPU loop (oracle_mlem_f32_omp) on the same arrays: `runs` runs of 1 warm-up iteration +
    k iterations, median.  k is the workload's 100 iterations, or more where 100 take less than budget_s / runs
    seconds (small matrices, to time a long enough run; one timed iteration sizes it).  Returns the median
    iterations/s, k, each run's seconds and the host description."""
    _, secs1 = O.mlem_f32_omp(A, AT, g, s32, f0, 1, threads)
    k = max(min_iters, int(budget_s / runs / max(secs1, 1e-6)))
    rates, secs_all = [], []
    for r in range(runs):
        O.mlem_f32_omp(A, AT, g, s32, f0, 1, threads)
        _, secs = O.mlem_f32_omp(A, AT, g, s32, f0, k, threads)
        rates.append(k / secs)
        secs_all.append(round(secs, 3))
        log(f"CPU baseline run {r + 1}/{runs}: {k} iterations in {secs:.1f} s ({k / secs:.2f} it/s)")
    nproc, phys = host_cpus()
    return float(np.median(rates)), k, secs_all, {
        "nproc": nproc, "physical_cores": phys, "cpu_model": cpu_model(),
        "omp": {v: os.environ.get(v) for v in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}}


def measured_traffic(workload_name, nnz, side, tile_width, tile_rows):
    """HBM bytes per launch of the blocked FP or BP kernel from profiles/traffic.json (rocprofv3 FETCH_SIZE and
    WRITE_SIZE in KiB per dispatch, each scaled by bytes/KiB read from the copy kernel), or None when it was not
    measured for this workload and tiling."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        if t.get("workload") != workload_name or t.get("nnz") != nnz or side not in t:
            return None
        if t.get("tile_width") != (tile_width or 1024) or t.get("tile_rows") != (tile_rows or 2048):
            return None
        cp = t["copy"]
        k_fetch = cp["bytes_read"] / (cp["fetch_kib"] * 1024.0)
        k_write = cp["bytes_written"] / (cp["write_kib"] * 1024.0)
        b = t[side]["fetch_kib"] * 1024.0 * k_fetch + t[side]["write_kib"] * 1024.0 * k_write
        return {"bytes": int(b), "source": f"profiles/traffic.json ({t.get('measured', '')})"}
    except (OSError, KeyError, ValueError, ZeroDivisionError):
        return None


def workload(name, threads):
    import mlem_amd as M
    if name == "c2":
        geo, counts, shapes, label = M.RING3D_1152(), 1e8, 8, "3D MLEM FP32, 2 rings x 576 crystals, 140x140x40 voxels"
    elif name == "mid":
        geo, counts, shapes, label = M.RING3D_576_MID(), 1e7, 6, "3D MLEM FP32, 2 rings x 288 crystals, 96x96x24 voxels"
    elif name == "c1":
        geo, counts, shapes, label = M.RING2D_288(), 1e6, 5, "2D MLEM FP32, 288-crystal ring, 64x64 voxels"
    elif name == "c3":
        geo, counts, shapes, label = (M.RING3D_1152_FINE(), 1e8, 8,
                                      "3D MLEM FP32, 2 rings x 576 crystals, 280x280x80 voxels, 128 sub-rays")
    else:
        raise SystemExit(f"unknown workload {name}")
    t = time.time()
    A = M.gen_system_matrix(geo, seed=1, threads=threads)
    log(f"system matrix {A.rows} x {A.cols}, nnz={A.nnz} ({A.nnz * 8 / 1e9:.2f} GB as f32+u32) in {time.time() - t:.1f} s")
    x_true = M.gen_phantom(geo, seed=1, n_shapes=shapes, a_min=0.5, a_max=4.0)
    g = M.gen_sinogram(A, x_true, counts, seed=1, threads=threads)
    return geo, A, g, label


VALUES = {"f32": 0, "bf16": 1, "fp16": 2, "uf16": 3}


def batch_sinograms(M, geo, A, B, threads):
    """B seeded phantoms' Poisson sinograms, counts log-uniform in 1e6-1e9."""
    rng = np.random.default_rng(2026)
    t = time.time()
    sinos = []
    for k in range(B):
        x = M.gen_phantom(geo, seed=1000 + k, n_shapes=4 + k % 5, a_min=0.5, a_max=4.0)
        sinos.append(M.gen_sinogram(A, x, float(10 ** rng.uniform(6, 9)), seed=2000 + k, threads=threads))
    log(f"{B} sinograms (counts log-uniform 1e6-1e9) in {time.time() - t:.1f} s")
    return sinos


def run_batch(args, M, geo, A, g0, m, s32, s64, threads, label, sinos=None, single=True):
    """Batched reconstructions: B seeded sinograms iterated together over one matrix; compared with the
    one-at-a-time FP32 path (the single Recon) timed in the same process (single=False: a matrix resident in
    the batches' layout only has no one-at-a-time path to time)."""
    B = args.batch
    if sinos is None:
        sinos = batch_sinograms(M, geo, A, B, threads)
    f0 = np.ones(A.cols, np.float32)
    t_single = t_plain = None

    def time_single(o):
        with M.Recon(m, g0, f0, sens=s32, options=o) as r:
            r.iterate(3)
            r.sync()
            t0 = time.perf_counter()
            r.iterate(20)
            r.sync()
            return (time.perf_counter() - t0) / 20

    if single:
        # one-at-a-time FP32: the fastest path (blocked kernels; a captured graph too when --graph), and the
        # reference's kernels (merge-path FP, CSR-vector BP on A^T)
        t_single = time_single(M.opts(spmv=0, using_trans=0, profile=0, use_graph=args.graph))
        o_plain = M.opts(spmv=0, using_trans=0, profile=0, use_graph=args.graph, tile_width=1)
        o_plain.bp_spmv = 2
        t_plain = time_single(o_plain)
    t = time.time()
    with M.Batch(m, B, values=VALUES[args.values], per_pass=args.per_pass,
                 tile_width=args.tile if args.tile > 1 else 0, tile_rows=args.tile_rows,
                 tile_groups=args.tile_groups, tile_threads=args.tile_threads, use_graph=args.graph) as b:
        for k in range(B):
            b.set(k, sinos[k], f0)
        info = b.info()
        log(f"batch set-up {time.time() - t:.2f} s: {info}")
        b.iterate(args.warmup)
        b.sync()
        M.device_sync()
        t0 = time.perf_counter()
        b.iterate(args.steps)
        b.sync()
        M.device_sync()
        elapsed = time.perf_counter() - t0
    # cpu_baseline leg: the FP64 OpenMP oracle reconstructs the first K sinograms on the host's physical cores
    # (timed: the CPU's recon/s), and, as the checker, its images are what every value type's batched images
    # are compared with, sinogram by sinogram.
    errs, cpu = {}, None
    K = min(B, args.batch_check) if (single and not args.no_cpu) else 0
    if K:
        import oracle as O
        AT = m.download_transpose()
        ct = baseline_threads()
        t = time.time()
        G, ref = np.stack(sinos[:K]), np.stack([f0] * K)
        for done in range(0, args.recon_iters, 10):  # chunks: the state between iterations is the images
            ref = O.mlem_f64_omp(A, AT, G, s64, ref, min(10, args.recon_iters - done), ct)
            log(f"FP64 oracle {min(done + 10, args.recon_iters)}/{args.recon_iters} iterations, {time.time() - t:.1f} s")
        secs = time.time() - t
        del AT
        nproc, phys = host_cpus()
        cpu = {"value": round(K / secs, 4), "unit": "recon/s", "cores": ct, "kind": "port",
               "sample": f"{K} sinograms x {args.recon_iters} FP64 MLEM iterations in one OpenMP pass per iteration "
                         f"(oracle_mlem_f64_omp, {secs:.1f} s) on {ct} threads, one per physical core of "
                         f"{cpu_model()}", "nproc": nproc, "physical_cores": phys}
        log(f"FP64 oracle: {K} sinograms x {args.recon_iters} iterations in {secs:.1f} s on {ct} threads")
        for vname in ("uf16", "fp16", "bf16", "f32"):
            with M.Batch(m, K, values=VALUES[vname], per_pass=args.per_pass,
                         tile_width=args.tile if args.tile > 1 else 0, tile_rows=args.tile_rows,
                         tile_groups=args.tile_groups, tile_threads=args.tile_threads) as b2:
                for k in range(K):
                    b2.set(k, sinos[k], f0)
                b2.iterate(args.recon_iters)
                b2.sync()
                e = [float(np.linalg.norm(b2.image(k) - ref[k]) / np.linalg.norm(ref[k])) for k in range(K)]
            errs[vname] = {"per_sinogram": [round(v, 8) for v in e], "max": round(max(e), 8),
                           "meets_1e-3": bool(max(e) <= 1e-3)}
            log(f"{vname} values, {args.per_pass} per pass: rel-L2 vs FP64 oracle after {args.recon_iters} "
                f"iterations " + ", ".join(f"{v:.2e}" for v in e))
    t_iter = elapsed / args.steps
    recon_s = B / (args.recon_iters * t_iter)
    single_recon_s = 1.0 / (args.recon_iters * t_single) if t_single else None
    log(f"batched: {args.steps} iterations of {B} sinograms in {elapsed * 1e3:.1f} ms = {t_iter * 1e3:.2f} ms per "
        f"iteration ({t_iter / B * 1e3:.3f} ms per sinogram-iteration); {recon_s:.2f} recon/s of "
        f"{args.recon_iters} iterations"
        + (f"; one-at-a-time FP32 {t_single * 1e3:.2f} ms, {recon_s / single_recon_s:.2f}x; with the reference's "
           f"kernels {t_plain * 1e3:.2f} ms, {t_plain / t_iter * B:.2f}x" if t_single else ""))
    out = {
        "metric": METRIC, "value": round(recon_s, 3), "unit": "recon/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t_iter * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": f"{args.values} values, f32 arithmetic", "data": "synthetic",
        "config": {"workload": label + f", batch of {B} sinograms", "batch": B, "per_pass": args.per_pass,
                   "values": args.values, "iterations_per_recon": args.recon_iters, "passes": info["passes"],
                   "graph": bool(args.graph), "lor_order": bool(args.lor_order and single),
                   "rows": A.rows, "cols": A.cols, "nnz": A.nnz},
        "one_at_a_time_f32": ({"ms_per_iteration": round(t_single * 1e3, 4), "recon_s": round(single_recon_s, 3)}
                              if t_single else None),
        "speedup_vs_one_at_a_time_f32": round(recon_s / single_recon_s, 3) if t_single else None,
        "one_at_a_time_f32_reference_kernels": (
            {"path": "merge-path FP + CSR-vector BP on A^T", "ms_per_iteration": round(t_plain * 1e3, 4),
             "speedup": round(t_plain / t_iter * B, 3)} if t_plain else None),
    }
    if errs:
        out["rel_l2_vs_fp64_oracle"] = {"iterations": args.recon_iters, "sinograms": K,
                                        "counts": [int(np.sum(sinos[k], dtype=np.int64)) for k in range(K)],
                                        "per_pass": args.per_pass, "by_values": errs}
    out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)
    m.close()


def serve_drive(srv, reqs, arrivals, mats):
    """Submits reqs[k] = (matrix index, sinogram index, iterations) at arrivals[k] seconds from the start and steps
    the server until every request is done.  Returns (latencies in s, makespan in s)."""
    order = np.argsort(arrivals, kind="stable")
    lat = np.zeros(len(reqs))
    live = {}
    nxt, done = 0, 0
    t0 = time.perf_counter()
    last = t0
    while done < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and arrivals[order[nxt]] <= now:
            k = int(order[nxt])
            mi, j, it = reqs[k]
            d = mats[mi]
            live[srv.submit(d["id"], d["sinos"][j], d["f0"], it)] = k
            nxt += 1
        wait = 0xFFFFFFFF
        if nxt < len(reqs):
            wait = max(0, int((arrivals[order[nxt]] - (time.perf_counter() - t0)) * 1e6))
            if not live:  # idle until the next arrival
                time.sleep(wait * 1e-6)
                continue
        srv.step(wait)
        for rid in list(live):
            k = live[rid]
            r = srv.poll(rid, mats[reqs[k][0]]["id"])
            if r is not None:
                lat[k] = r[1]
                del live[rid]
                done += 1
                last = time.perf_counter()
    return lat, last - t0


def serial_fifo(M, reqs, arrivals, mats):
    """The same requests served one at a time, first come first served, by the single-reconstruction FP32 path
    (a reconstruction handle per request: upload, iterations, image back).  Returns (latencies, makespan) in s."""
    lat = np.zeros(len(reqs))
    o = M.opts(spmv=0, using_trans=0)
    t0 = time.perf_counter()
    for k in np.argsort(arrivals, kind="stable"):
        now = time.perf_counter() - t0
        if arrivals[k] > now:
            time.sleep(arrivals[k] - now)
        mi, j, it = reqs[k]
        d = mats[mi]
        with M.Recon(d["m"], d["sinos"][j], d["f0"], sens=d["s32"], options=o) as r:
            r.iterate(it)
            r.sync()
            r.image()
        lat[k] = time.perf_counter() - t0 - arrivals[k]
    return lat, time.perf_counter() - t0


def set_lor_order(M, m, A, geo, threads):
    """The batches' row order (mlem_matrix_set_row_order) from each LOR's transaxial line on the image grid."""
    t = time.time()
    m.set_row_order(M.lor_order(A, tuple(geo.grid), threads=threads))
    log(f"row order of the batched layouts (LORs by transaxial line) {time.time() - t:.2f} s")


def run_serve(args, M, threads):
    """Serving mode: reconstruction requests over several resident matrices, continuous batching
    (mlem_server_*).  Phase 1 submits every request at once (throughput, recon/s); phase 2 replays the mix with
    Poisson arrivals at a fraction of that rate (latency p50/p99).  Both are compared with the one-at-a-time
    FP32 path serving the same requests first come first served (measured in the same process)."""
    names = args.serve.split(",")
    mats = []
    for name in names:
        geo, A, g, label = workload(name, threads)
        m = M.ResidentMatrix(A)
        m.build_transpose()
        s32, _ = m.sensitivity()
        if args.lor_order:
            set_lor_order(M, m, A, geo, threads)
        rng = np.random.default_rng(len(mats) + 5)
        counts = float(np.sum(g, dtype=np.float64))
        sinos = [g] + [M.gen_sinogram(A, M.gen_phantom(geo, seed=70 + k, n_shapes=4 + k, a_min=0.5, a_max=4.0),
                                      counts * float(rng.uniform(0.3, 3.0)), seed=80 + k, threads=threads)
                       for k in range(3)]
        f0 = np.ones(A.cols, np.float32)
        with M.Recon(m, g, f0, sens=s32, options=M.opts(spmv=0, using_trans=0)) as r:
            r.iterate(3)
            r.sync()
            t0 = time.perf_counter()
            r.iterate(20)
            r.sync()
            t_single = (time.perf_counter() - t0) / 20
        log(f"{name}: {label}; one-at-a-time FP32 {t_single * 1e3:.3f} ms per iteration")
        mats.append(dict(name=name, label=label, m=m, sinos=sinos, f0=f0, t_single=t_single, s32=s32, rows=A.rows,
                         cols=A.cols, nnz=A.nnz))
        del A
    srv = M.Server(values=VALUES[args.values], per_pass=args.per_pass, use_graph=args.graph)
    t = time.time()
    for d in mats:
        d["id"] = srv.add_matrix(d["m"], slots=args.serve_slots)
    log(f"server set-up {time.time() - t:.2f} s: {srv.info()}")
    rng = np.random.default_rng(2027)
    n = args.serve_requests
    if "-" in args.serve_iters:  # uniform integers lo..hi
        lo, hi = (int(c) for c in args.serve_iters.split("-"))
        draw_iters = lambda: int(rng.integers(lo, hi + 1))
        iters_desc = f"uniform {lo}..{hi}"
    else:
        choices = [int(c) for c in args.serve_iters.split(",")]
        draw_iters = lambda: int(rng.choice(choices))
        iters_desc = choices
    reqs = [(int(rng.integers(len(mats))), int(rng.integers(4)), draw_iters()) for _ in range(n)]
    # warm-up: one request per matrix
    serve_drive(srv, [(i, 0, 2) for i in range(len(mats))], np.zeros(len(mats)), mats)
    lat1, span1 = serve_drive(srv, reqs, np.zeros(n), mats)
    recon_s = n / span1
    ser1, sspan1 = serial_fifo(M, reqs, np.zeros(n), mats)
    serial_s = n / sspan1
    log(f"backlog of {n} requests served in {span1:.3f} s: {recon_s:.2f} recon/s (one at a time: {sspan1:.3f} s, "
        f"{serial_s:.2f} recon/s; {recon_s / serial_s:.2f}x)")
    pct = lambda a, q: round(float(np.percentile(a, q)) * 1e3, 3)
    # Poisson arrivals over a sweep of rates (fractions of the backlog throughput); the same arrival times are
    # replayed through the one-at-a-time path.  The first load is the headline latency.
    sweep = []
    for load in [float(c) for c in args.serve_load.split(",")]:
        rate = load * recon_s
        arr = np.cumsum(np.random.default_rng(2028).exponential(1.0 / rate, n))
        lat, _ = serve_drive(srv, reqs, arr, mats)
        ser, _ = serial_fifo(M, reqs, arr, mats)
        log(f"Poisson arrivals at {rate:.2f} req/s ({load} x): p50 {pct(lat, 50)} ms, p99 {pct(lat, 99)} ms "
            f"(one at a time: p50 {pct(ser, 50)} ms, p99 {pct(ser, 99)} ms)")
        sweep.append(dict(load=load, rate=round(rate, 3), p50=pct(lat, 50), p99=pct(lat, 99), max=pct(lat, 100),
                          one_at_a_time_p50=pct(ser, 50), one_at_a_time_p99=pct(ser, 99)))
    head = sweep[0]
    info = srv.info()
    stats = srv.stats()
    log(f"passes over a matrix started: {stats['pass_sweeps']}; requests moved to a lower pass to empty one: "
        f"{stats['slot_moves']}")
    srv.close()
    for d in mats:
        d["m"].close()
    out = {
        "metric": METRIC, "value": round(recon_s, 3), "unit": "recon/s", "n_gpus": 1, "steps": n,
        "warmup": len(mats), "ms_per_step": round(span1 / n * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": f"{args.values} values, f32 arithmetic", "data": "synthetic",
        "config": {"workload": "served mix: " + " + ".join(d["label"] for d in mats), "matrices": names,
                   "requests": n, "iterations_per_request": iters_desc, "slots_per_matrix": args.serve_slots,
                   "per_pass": args.per_pass, "values": args.values, "graph": bool(args.graph),
                   "lor_order": bool(args.lor_order), "hbm_used_bytes": info["used"],
                   "pass_sweeps": stats["pass_sweeps"], "slot_moves": stats["slot_moves"]},
        "latency_ms": {"arrivals": f"Poisson, {head['rate']:.3f} req/s ({head['load']} x backlog throughput)",
                       "p50": head["p50"], "p99": head["p99"], "max": head["max"],
                       "backlog_p50": pct(lat1, 50), "backlog_p99": pct(lat1, 99)},
        "latency_sweep_ms": sweep,
        "one_at_a_time_f32": {"recon_s": round(serial_s, 3), "p50_ms": head["one_at_a_time_p50"],
                              "p99_ms": head["one_at_a_time_p99"],
                              "ms_per_iteration": {d["name"]: round(d["t_single"] * 1e3, 4) for d in mats},
                              "note": "the same requests and arrivals served first come first served by the "
                                      "single-reconstruction FP32 path, one handle per request, measured"},
        "speedup_vs_one_at_a_time_f32": round(recon_s / serial_s, 3),
    }
    print(json.dumps(out), flush=True)


def parse_args(argv=None):
    """The command line, with the defaults that depend on the workload resolved: 16-bit (UF16) values, graphs and
    the autotuner for batches and serving; graphs, 2000 timed iterations and CSR-vector BP (the workload's own
    pairing with merge-based FP) for the launch-bound 2D ring."""
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed MLEM iterations (default: 100; 2000 on the 2D ring, whose iterations take about 25 us)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c2", choices=["c2", "mid", "c1", "c3"],
                    help="c3: the 100 GB-class matrix, resident as the tiles of A and A^T only (no CPU baseline: "
                         "A and A^T do not fit the host memory of one command)")
    ap.add_argument("--c3-csr", action="store_true",
                    help="c3: keep A and A^T resident as CSR instead (plain merge / CSR-vector kernels)")
    ap.add_argument("--fp", default="merge", choices=["merge", "vector"])
    ap.add_argument("--bp", default=None, choices=["merge", "vector"],
                    help="default: vector for the 2D ring (merge-based FP with CSR-vector BP on A^T is the pairing "
                         "that workload names), else merge")
    ap.add_argument("--secsize", type=int, default=8)
    ap.add_argument("--tile", type=int, default=0, help="merge-path tiling: 0 auto, 1 off, else segment width")
    ap.add_argument("--tile-rows", type=int, default=0, help="rows per tile block (0: default)")
    ap.add_argument("--tile-groups", type=int, default=0, help="workgroups per blocked SpMV (0: default)")
    ap.add_argument("--tile-threads", type=int, default=0,
                    help="work-items per workgroup of the kernel for 4 or 8 per pass: 128, 256, 512 (0: default)")
    ap.add_argument("--graph", type=int, default=None,
                    help="replay captured hipGraphs (default: 1 for batches and serving, whose workloads ask for "
                         "graph-captured iterations, and for the launch-bound 2D ring, else 0)")
    ap.add_argument("--autotune", type=int, nargs="?", const=1, default=None,
                    help="time the candidate tile shapes and workgroup counts first (mlem_autotune) and use the best "
                         "(default: 1 for batches, whose workload asks for autotuned shapes, else 0)")
    ap.add_argument("--threads", type=int, default=0, help="host threads (0: all)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (profiling runs)")
    ap.add_argument("--batch", type=int, default=0,
                    help="batched mode: B sinograms (seeded phantoms, counts log-uniform 1e6-1e9) over the matrix")
    ap.add_argument("--per-pass", type=int, default=4, help="batched mode: sinograms per pass over the matrix")
    ap.add_argument("--values", default=None, choices=["f32", "bf16", "fp16", "uf16"],
                    help="stored matrix values (default: f32 for one reconstruction, uf16 for --batch and --serve)")
    ap.add_argument("--recon-iters", type=int, default=100, help="iterations per reconstruction (recon/s)")
    ap.add_argument("--lor-order", type=int, default=1,
                    help="batched and serving modes: 1 = the batches' tiles take A's rows (LORs) in the order of "
                         "mlem_lor_order (fewer padded entries in A^T's tiles), 0 = as stored")
    ap.add_argument("--batch-check", type=int, default=8,
                    help="batched mode: sinograms reconstructed by the FP64 oracle (CPU baseline and checker) and "
                         "compared, for every value type, with their batched images")
    ap.add_argument("--serve", default="",
                    help="serving mode over these workloads (comma list of c1, mid, c2): continuous batching")
    ap.add_argument("--serve-requests", type=int, default=48, help="serving mode: number of requests")
    ap.add_argument("--serve-slots", type=int, default=16, help="serving mode: slots per matrix")
    ap.add_argument("--serve-iters", default="10-100",
                    help="serving mode: iterations per request, uniform over lo-hi or drawn from a comma list")
    ap.add_argument("--serve-load", default="0.7,0.3,0.5,0.9",
                    help="serving mode: Poisson arrival rates as fractions of the backlog throughput (the first is "
                         "the headline latency)")
    ap.add_argument("--load-test", action="store_true",
                    help="also write the matrix as a csr4 file and time the streamed loader on it")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB device-to-device copy (HBM counter calibration)")
    args = ap.parse_args(argv)
    if args.values is None:
        args.values = "uf16" if (args.batch or args.serve) else "f32"
    if args.graph is None:
        args.graph = 1 if (args.serve or args.batch or args.workload == "c1") else 0
    if args.steps is None:
        args.steps = 2000 if args.workload == "c1" and not (args.serve or args.batch) else 100
    if args.autotune is None:
        args.autotune = 1 if args.batch else 0
    if args.bp is None:
        args.bp = "vector" if args.workload == "c1" and not (args.batch or args.serve) else "merge"
    return args


def main():
    args = parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(f"bench.py: this engine keeps the whole system matrix resident on ONE GPU; "
              f"--gpus {max(args.gpus, world)} is not supported (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import mlem_amd as M

    threads = args.threads or min(16, os.cpu_count() or 1)
    name, mem, cus = M.device_info(0)
    log(f"device {name}, {mem / 2**30:.0f} GiB, {cus} CUs; host {cpu_model()}, {threads} threads")
    if args.serve:
        return run_serve(args, M, threads)
    geo, A, g, label = workload(args.workload, threads)

    t = time.time()
    tiles_only = args.workload == "c3" and not args.c3_csr
    batch_sinos = None
    if tiles_only:
        m = M.ResidentMatrix.tiles_only(A, tile_width=args.tile if args.tile > 1 else 0, tile_rows=args.tile_rows,
                                        values=VALUES[args.values], per_pass=args.per_pass if args.batch else 1)
        log(f"tiles of A and A^T built without resident CSR copies (A^T's from column ranges of A on the device, "
            f"A's from row chunks of the host CSR) {time.time() - t:.2f} s; "
            f"{m.info()['device_bytes'] / 1e9:.1f} GB resident")
    else:
        m = M.ResidentMatrix(A)
        log(f"upload A {time.time() - t:.2f} s")
    if tiles_only and args.batch:  # the sinograms need the host copy
        batch_sinos = batch_sinograms(M, geo, A, args.batch, threads)
    if args.workload == "c3":  # the host copy is not needed any more (and is ~120 GB)
        class _Dims:
            pass
        dims = _Dims()
        dims.rows, dims.cols, dims.nnz = A.rows, A.cols, A.nnz
        A = dims
        args.no_cpu = True
        if not tiles_only:
            args.tile = 1  # the tiles do not fit in HBM beside both CSR copies: plain merge-path kernels
        log("host copy of A released; CPU baseline skipped (A and A^T do not fit the host memory of one command)")
    if not tiles_only:
        t = time.time()
        ranges = m.build_transpose()
        log(f"device transpose (stable counting sort{f', {ranges} column ranges' if ranges else ''}) "
            f"{time.time() - t:.2f} s")
    t = time.time()
    s32, s64 = m.sensitivity()
    log(f"FP64 sensitivity {time.time() - t:.2f} s")
    f0, init = M.initial_image(g, A.rows, s32)
    if args.batch and args.lor_order and not tiles_only:
        set_lor_order(M, m, A, geo, threads)
    if args.autotune and args.tile != 1 and not tiles_only:
        t = time.time()
        tune = m.autotune(per_pass=args.per_pass if args.batch else 1,
                          values=VALUES[args.values] if args.batch else 0)
        args.tile, args.tile_rows, args.tile_groups = tune["tile_width"], tune["tile_rows"], tune["tile_groups"]
        args.tile_threads = tune["tile_threads"]
        log(f"autotune {time.time() - t:.1f} s over {tune['candidates']} candidates: tiles {args.tile} x "
            f"{args.tile_rows}, {args.tile_groups} workgroups of {args.tile_threads or 'default'} work-items "
            f"(FP {tune['fp_ms']:.3f} ms, BP {tune['bp_ms']:.3f} ms)")

    if args.batch:
        return run_batch(args, M, geo, A, g, m, s32, s64, threads, label, sinos=batch_sinos, single=not tiles_only)

    # FP and BP kernels chosen independently: spmv selects FP, bp_* selects BP.
    fp_merge = args.fp == "merge"
    o = M.opts(spmv=0 if fp_merge else 1, using_trans=0, secsize_fp=args.secsize, secsize_bp=args.secsize,
               profile=0 if args.graph else 1, use_graph=args.graph, tile_width=args.tile,
               tile_rows=args.tile_rows, tile_groups=args.tile_groups, values=VALUES[args.values])
    o.bp_spmv = 1 if args.bp == "merge" else 2
    t = time.time()
    with M.Recon(m, g, f0, sens=s32, options=o) as r:
        log(f"reconstruction set-up (tiles of A and A^T, work lists) {time.time() - t:.2f} s")
        r.iterate(args.warmup)
        r.sync()
        ll0 = r.loglik() if args.workload == "c3" else None
        r.reset_times()
        M.device_sync()
        t0 = time.perf_counter()
        r.iterate(args.steps)
        r.sync()
        M.device_sync()
        t1 = time.perf_counter()
        phase, n_timed = r.phase_times()
        img = r.image()
        if ll0 is not None:  # MLEM does not decrease the Poisson log-likelihood (a size-independent check)
            ll1 = r.loglik()
            log(f"log-likelihood after {args.warmup} iterations {ll0:.9e}, after {args.warmup + args.steps}: "
                f"{ll1:.9e} ({'increased' if ll1 >= ll0 else 'DECREASED'})")
    elapsed = t1 - t0
    its = args.steps / elapsed
    rel_vs_f32 = None
    if args.values != "f32" and not tiles_only:  # the same iterations with f32 values: the 16-bit image error
        o32 = M.opts(spmv=0, using_trans=0, profile=0, tile_width=args.tile, tile_rows=args.tile_rows,
                     tile_groups=args.tile_groups)
        with M.Recon(m, g, f0, sens=s32, options=o32) as r32:
            r32.iterate(args.warmup + args.steps)
            ref32 = r32.image()
        rel_vs_f32 = float(np.linalg.norm(img.astype(np.float64) - ref32) / np.linalg.norm(ref32.astype(np.float64)))
        log(f"{args.values} values: rel-L2 vs the f32-value image after {args.warmup + args.steps} iterations "
            f"{rel_vs_f32:.2e}")
    rows, cols, nnz = A.rows, A.cols, A.nnz
    vb = 4 if args.values == "f32" else 2  # bytes per stored value (16-bit values add a 4-byte scale per row of A)
    b_fp = (vb + 4) * nnz + 8 * (rows + 1) + 4 * cols + 8 * rows + (4 * rows if vb == 2 else 0)
    b_bp = (vb + 4) * nnz + 8 * (cols + 1) + 4 * rows + 12 * cols  # A^T, rowptr, ratio, f r/w, sens
    b_iter = b_fp + b_bp
    log(f"{args.steps} iterations in {elapsed * 1e3:.1f} ms: {its:.1f} it/s, {b_iter * its / 1e9:.0f} GB/s; "
        f"sum f = {float(np.sum(img, dtype=np.float32)):.6g}")

    roofline = None
    blocked = args.tile != 1
    if n_timed:
        ph = [v / n_timed for v in phase]  # FP kernel, FP epilogue, BP kernel, BP epilogue (ms)
        k_fp = (vb + 4) * nnz + 8 * (rows + 1) + 4 * cols
        k_bp = (vb + 4) * nnz + 8 * (cols + 1) + 4 * rows
        gbs = lambda b, ms: b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        log(f"FP kernel {ph[0]:.3f} ms ({gbs(k_fp, ph[0]):.0f} GB/s) + epilogue {ph[1]:.3f} ms; "
            f"BP kernel {ph[2]:.3f} ms ({gbs(k_bp, ph[2]):.0f} GB/s) + epilogue {ph[3]:.3f} ms")
        dom_fp = ph[0] >= ph[2]
        side = "fp" if dom_fp else "bp"
        ms = ph[0] if dom_fp else ph[2]
        kb = k_fp if dom_fp else k_bp
        ach = gbs(kb, ms)
        algo = args.fp if dom_fp else args.bp
        if algo == "merge" and blocked:
            kname = f"blocked_spmv_kernel (LDS-staged x segments, {vb + 2}-byte entries)"
        elif algo == "merge":
            kname = "merge_spmv_kernel"
        else:
            kname = "csr_vector_kernel"
        roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": f"{'FP' if dom_fp else 'BP'} {kname}",
                    "algorithmic_bytes_per_launch": kb, "avg_launch_ms": round(ms, 4),
                    "phases_ms": {"fp_kernel": round(ph[0], 4), "fp_epilogue": round(ph[1], 4),
                                  "bp_kernel": round(ph[2], 4), "bp_epilogue": round(ph[3], 4)}}
        tr = (measured_traffic(args.workload, nnz, side, args.tile, args.tile_rows)
              if (algo == "merge" and blocked and not args.tile_groups and args.values == "f32"
                  and os.environ.get("MLEM_TILE_ORDER", "") != "cols") else None)
        if tr:
            roofline["traffic"] = tr["bytes"]
            roofline["traffic_source"] = tr["source"]
            roofline["drawn_gbs"] = round(gbs(tr["bytes"], ms), 1)
            roofline["drawn_frac"] = round(gbs(tr["bytes"], ms) / HBM_PEAK_GBS, 4)
            log(f"measured HBM traffic per {side.upper()} launch: {tr['bytes'] / 1e9:.2f} GB "
                f"({tr['bytes'] / (ms * 1e-3) / 1e9:.0f} GB/s drawn)")

    if args.calibrate_copy:
        n = 1 << 28
        a = M.DeviceArray(n, np.float32)
        b = M.DeviceArray(n, np.float32)
        M.fill_f32(a, 1.0)
        M.device_sync()
        t = time.perf_counter()
        for _ in range(3):
            M.copy_f32(b, a)
        M.device_sync()
        log(f"copy calibration: 3 x 1 GiB copy_kernel, {3 * 2 * a.nbytes / (time.perf_counter() - t) / 1e9:.0f} GB/s")

    load = None
    if args.load_test:
        import tempfile
        d = tempfile.mkdtemp(prefix="mlem_load_", dir=os.environ.get("TMPDIR", "/tmp"))
        path = os.path.join(d, "A.csr4")
        t = time.time()
        M.write_csr4(path, A)
        log(f"wrote {path} ({os.path.getsize(path) / 1e9:.2f} GB) in {time.time() - t:.1f} s")
        try:
            for rep in range(2):
                lm, st = M.ResidentMatrix.load_csr4(path, threads=threads)
                lm.close()
                log(f"streamed csr4 load: {st['bytes'] / 1e9:.2f} GB in {st['stream_s']:.3f} s "
                    f"({st['bytes'] / 1e9 / st['stream_s']:.1f} GB/s), open {st['open_s'] * 1e3:.1f} ms")
            t = time.time()
            H = M.read_csr4(path, threads=threads)
            t_read = time.time() - t
            t = time.time()
            hm = M.ResidentMatrix(H)
            t_up = time.time() - t
            hm.close()
            log(f"host path: read_csr4 + CSR build {t_read:.2f} s, upload {t_up:.2f} s")
            load = {"bytes": st["bytes"], "stream_s": round(st["stream_s"], 4),
                    "gbs": round(st["bytes"] / 1e9 / st["stream_s"], 2), "threads": threads,
                    "host_path_s": round(t_read + t_up, 3), "file": "warm page cache (just written)"}
        finally:
            os.remove(path)
            os.rmdir(d)

    cpu = witness = None
    if not args.no_cpu:
        # cpu_baseline leg: the FP32 OpenMP CPU loop timed on the host's physical cores (1 warm-up + the
        # workload's iteration count, median of 3), and, as the checker, the FP64 OpenMP oracle over the same
        # iterations against a GPU reconstruction of the timed configuration.
        import oracle as O
        AT = m.download_transpose()
        ct = baseline_threads()
        rate, k, run_s, host = cpu_baseline_f32(O, A, AT, g, s32, f0, ct)
        cpu = {"value": round(rate, 4), "unit": "iters/s", "cores": ct, "kind": "port",
               "sample": f"FP32 MLEM of the same workload, 1 warm-up + {k} iterations, median of {len(run_s)} runs "
                         f"({', '.join(f'{v:.1f}' for v in run_s)} s) on {ct} threads, one per physical core "
                         f"({host['cpu_model']}); OpenMP row-parallel FP on A and BP on A^T (oracle_mlem_f32_omp)",
               **host}
        log(f"CPU baseline {cpu['value']} it/s on {ct} threads ({host['physical_cores']} physical cores of "
            f"{host['nproc']} CPUs; runs {run_s} s)")
        if args.workload == "c1":  # the small matrix on one core as well (same protocol, a shorter budget)
            rate1, k1, run_s1, _ = cpu_baseline_f32(O, A, AT, g, s32, f0, 1, budget_s=9.0)
            cpu["one_core"] = {"value": round(rate1, 4), "unit": "iters/s", "iterations": k1, "runs_s": run_s1}
            log(f"CPU baseline on one core: {cpu['one_core']['value']} it/s (runs {run_s1} s)")
        if args.values == "f32":
            o_w = M.opts(spmv=0 if fp_merge else 1, using_trans=0, secsize_fp=args.secsize, secsize_bp=args.secsize,
                         use_graph=args.graph, tile_width=args.tile, tile_rows=args.tile_rows,
                         tile_groups=args.tile_groups)
            o_w.bp_spmv = o.bp_spmv
            img_w = M.mlem_run(m, g, f0, args.recon_iters, sens=s32, options=o_w)
            t = time.time()
            ref = f0
            for done in range(0, args.recon_iters, 10):  # chunks: the state between iterations is the image
                ref = O.mlem_f64_omp(A, AT, g, s64, ref, min(10, args.recon_iters - done), ct)
                log(f"witness: FP64 oracle {min(done + 10, args.recon_iters)}/{args.recon_iters} iterations, "
                    f"{time.time() - t:.1f} s")
            err = float(np.linalg.norm(img_w - ref) / np.linalg.norm(ref))
            witness = {"iterations": args.recon_iters, "rel_l2_vs_fp64_oracle": err, "bar": 1e-5,
                       "oracle": "oracle_mlem_f64_omp (FP64, FP64 sensitivity; bitwise the serial oracle)",
                       "oracle_s": round(time.time() - t, 1)}
            log(f"witness: the GPU image after {args.recon_iters} iterations is {err:.2e} rel-L2 from the FP64 "
                f"oracle ({witness['oracle_s']} s)")
        del AT
    m.close()

    out = {
        "metric": METRIC, "value": round(its, 3), "unit": "iters/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed * 1e3 / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.values == "f32" else f"{args.values} values, f32 arithmetic", "data": "synthetic",
        "algorithmic_gbs": round(b_iter * its / 1e9, 1),
        "config": {"workload": label, "rows": rows, "cols": cols, "nnz": nnz,
                   "matrix_gb": round(8 * nnz / 1e9, 2), "bytes_per_iteration": b_iter,
                   "fp": args.fp + (" blocked" if args.fp == "merge" and blocked else ""),
                   "bp": args.bp + (" blocked" if args.bp == "merge" and blocked else "") + " on A^T",
                   "tile_width": args.tile, "tile_rows": args.tile_rows, "tile_groups": args.tile_groups,
                   "graph": bool(args.graph),
                   "parallelism": "single GPU, " + ("tiles of A and A^T resident (no CSR copies)" if tiles_only
                                                    else "matrix and transpose resident")},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    if witness:
        out["witness"] = witness
    if rel_vs_f32 is not None:
        out["rel_l2_vs_f32_values"] = {"iterations": args.warmup + args.steps, "value": rel_vs_f32}
    if load:
        out["load_csr4"] = load
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
