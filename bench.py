"""Benchmark of the TRMM hot path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--n N] [--m M] [--bands P]
                    [--map reg|balanced] [--shape boxes|trapezoid] [--no-cpu]

A step is one C = tril(A) * B (FP64) over the whole matrix, with A, B and C
resident in HBM before the timed region.  Default workload: N = 32768,
M = 8192, balanced mapping over 16 bands, trapezoid tile walk, inputs
U[-1,1) with exponents spread over 2^[-8,8] (seeded generator, on device).

Prints one JSON line.  value = M*N^2 / (mean step time) in TFLOP/s.
roofline: the dominant kernel (the tile kernel) against FP64 MFMA peak,
timed with HIP events on the stream the kernel runs on.
Beside it: the split-precision leg on the same inputs (time, and the error
statement of each call), the 512-problem batch (routed by measured speed,
and with its split half forced onto the int8 path), the reference's other
two configurations, and the served stream (large requests first come first
served; then the same stream with them shortest first, which lowers p99 and
the mean while the largest requests wait longer: p999 and max beside it).
cpu_baseline: the reference's per-rank stage loop (cblas_dgemm + cblas_dtrmm
with MKL, sequential, one thread per rank) replayed on the host cores for a
bounded column sample of the same matrices, median of 5 runs, for this
workload and the reference's N=2048 and N=16384 configurations.  The same
leg computes sampled rows of the exact product (double-double sums) and
reports the error of the FP64 and split results on them, and times the
ScaLAPACK pdtrmm comparator when its build is present.
vendor_gemm: the vendor int8 and FP64 GEMMs (torch: hipBLASLt / rocBLAS)
timed on the same GPU in a child process, as practical ceilings
(--no-vendor-ceiling skips them).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from uvatrmm_amd import workloads  # noqa: E402  (seeded workload definitions)

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix peak
# HBM bytes per launch of the FP64 tile kernel (256x128 tiles) at N=32768,
# M=8192 (balanced, trapezoid, 16 bands), from two rocprofv3 counter runs of
# this script: 2.0 x FETCH_SIZE (58,506,032 KiB; the factor is what
# FETCH_SIZE reads for k_row_abs_sums streaming exactly 2 GiB in 128-byte
# requests, the request size of the tile kernel's 128-byte row segments) +
# WRITE_SIZE (2,097,152 KiB, exactly C).  Kept per workload; other workloads
# report null.
MEASURED_TRAFFIC = {(32768, 8192, 16, "balanced", "trapezoid"): 2.0 * 58506032 * 1024 + 2097152 * 1024}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


CPU_RUNS = 5


def host_ranks():
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(avail, int(os.environ.get("OMP_NUM_THREADS", avail)), 16))


def stage_loop_median(orc, n, m_total, bands, mapping, ranks, w, seed, spread, have_mkl):
    """The reference stage loop on `ranks` host threads (one per rank, each
    over its own w columns with sequential MKL), CPU_RUNS times; returns
    (median seconds, all runs, C of the last run, the flops of one run)."""
    A = np.empty((n, n))
    orc.L.orc_fill_seeded_A(n, seed, spread, orc._p(A), n)
    B = np.empty((n, m_total))
    orc.L.orc_fill_seeded_B(n, m_total, seed, spread, orc._p(B), m_total)
    Bs = np.ascontiguousarray(B[:, : ranks * w])
    del B
    runs, Cs = [], None
    for _ in range(CPU_RUNS):
        Cs = np.zeros((n, ranks * w))
        secs = orc.L.orc_ranks_stage_loop(n, w, bands, 1 if mapping == "balanced" else 0, ranks,
                                          orc._p(A), n, orc._p(Bs), ranks * w, orc._p(Cs),
                                          ranks * w, 1 if have_mkl else 0)
        if secs < 0:
            raise RuntimeError("stage loop failed")
        runs.append(secs)
    return float(np.median(runs)), runs, Cs, float(n) * n * (ranks * w)


def scalapack_leg(args, C_gpu_cols, ranks):
    """The reference's second comparator, MKL ScaLAPACK pdtrmm on a 2-D
    block-cyclic grid (nb = 256), through oracle/scalapack_trmm.x under
    MPICH: this workload on a 1024-column sample of B (its norm checked
    against the GPU's same columns), and the N=2048 (all columns) and
    N=16384 (2048 columns) configurations; median of CPU_RUNS runs each."""
    import subprocess

    exe = os.path.join(ROOT, "oracle", "scalapack_trmm.x")
    mpiexec = "/opt/conda/bin/mpiexec"
    if not (os.path.exists(exe) and os.path.exists(mpiexec)):
        return {"skipped": "scalapack_trmm.x or mpiexec missing"}
    grids = {16: (4, 4), 8: (2, 4), 4: (2, 2), 2: (1, 2), 1: (1, 1)}
    p = max(k for k in grids if k <= ranks)
    out = []
    for n, m, cmp_gpu in ((args.n, 1024, True), (2048, 2048, False), (16384, 2048, False)):
        pr, pc = grids[p]
        # the sample's B is the first m columns of the workload's own B
        cmd = [mpiexec, "-n", str(p), exe, str(n), str(m), "256", str(pr), str(pc), str(CPU_RUNS),
               str(args.seed), str(args.spread), str(args.m if cmp_gpu else m)]
        env = dict(os.environ, OMP_NUM_THREADS="1", MKL_NUM_THREADS="1")
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600, env=env)
        if r.returncode != 0:
            out.append({"workload": f"N={n} M={m}", "failed": (r.stderr or r.stdout)[-400:]})
            continue
        f = r.stdout.split()
        med, fro = float(f[6]), float(f[9])
        rec = {"workload": f"pdtrmm L,L,N,N N={n} M={m}", "value": round(float(m) * n * n / med * 1e-12, 4),
               "unit": "TFLOP/s", "cores": p, "grid": f"{pr}x{pc}", "nb": 256, "kind": "port",
               "seconds": med, "min_s": float(f[7]), "max_s": float(f[8]), "runs": CPU_RUNS}
        if cmp_gpu and C_gpu_cols is not None and C_gpu_cols.shape[1] >= m:
            g = float(np.linalg.norm(C_gpu_cols[:, :m]))
            rec["gpu_vs_scalapack_norm_rel_diff"] = abs(g - fro) / fro
        out.append(rec)
        log(f"scalapack pdtrmm N={n} M={m}: {med:.3f} s, {rec['value']} TFLOP/s on {p} ranks")
    return out


def cpu_baseline(args, C_gpu_cols, rows_gpu):
    """Reference stage loop on host cores (MKL) for column samples of this
    workload and of the reference's N=2048 / N=16384 configurations, and the
    exact sampled rows that check the GPU results."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import Oracle  # the oracle library: CPU comparator and checker

    orc = Oracle()
    ver = ctypes.create_string_buffer(256)
    have_mkl = bool(orc.L.orc_blas_available(ver, 256))
    ranks = host_ranks()
    n, m = args.n, args.m
    w = args.cpu_cols_per_rank if have_mkl else max(2, args.cpu_cols_per_rank // 8)
    w = min(w, m // ranks)
    lib = f"MKL {ver.value.decode()[:40]} sequential" if have_mkl else "C loops"
    log(f"cpu baseline: {ranks} ranks x {w} columns, mkl={have_mkl}, median of {CPU_RUNS}")
    med, runs, Cs, flops = stage_loop_median(orc, n, m, args.bands, args.map, ranks, w, args.seed,
                                             args.spread, have_mkl)
    err = None
    if C_gpu_cols is not None:
        G = C_gpu_cols[:, : ranks * w]
        err = float(np.linalg.norm(G - Cs) / np.linalg.norm(Cs))
    del Cs
    out = {
        "value": flops / med * 1e-12,
        "unit": "TFLOP/s",
        "cores": ranks,
        "kind": "port",
        "sample": (f"reference stage loop (dgemm box + dtrmm triangle per band, "
                   f"{args.map} {args.bands} bands) for {ranks} ranks x {w} columns of the "
                   f"N={n} workload ({ranks * w} of {m} columns), {lib}, median of {CPU_RUNS} "
                   f"runs of {med:.2f} s"),
        "seconds": med,
        "runs_s": [round(r, 3) for r in runs],
        "gpu_vs_cpu_rel_fro": err,
    }
    # the reference's other two configurations: N=2048 reg P=4 (4 ranks, each
    # its 512 columns, the whole product) and N=16384 balanced P=16 (16 ranks,
    # a 128-column sample of each rank's 1024)
    other = []
    for cn, cp, cmap, cw in ((2048, 4, "reg", 512), (16384, 16, "balanced", 128)):
        cr = min(cp, ranks)
        cmed, cruns, _, cfl = stage_loop_median(orc, cn, cn, cp, cmap, cr, cw, args.seed,
                                                args.spread, have_mkl)
        other.append({"workload": f"N=M={cn} {cmap} P={cp}", "value": round(cfl / cmed * 1e-12, 4),
                      "unit": "TFLOP/s", "cores": cr, "kind": "port",
                      "sample": f"{cr} ranks x {cw} columns ({cr * cw} of {cn}), {lib}",
                      "seconds": round(cmed, 4), "runs_s": [round(r, 4) for r in cruns]})
        log(f"cpu baseline N={cn}: {cfl / cmed * 1e-12:.4f} TFLOP/s on {cr} cores")
    out["other_configs"] = other
    out["scalapack"] = scalapack_leg(args, C_gpu_cols, ranks)
    # exact rows (double-double sums of the seeded inputs, generated on the fly)
    rows = np.asarray(rows_gpu["rows"], dtype=np.int64)
    t0 = time.time()
    ref = np.empty((len(rows), m))
    orc.L.orc_seeded_rows_dd(n, m, args.seed, args.spread,
                             rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(rows),
                             orc._p(ref), m)
    errs = {k: float(np.linalg.norm(v - ref) / np.linalg.norm(ref))
            for k, v in rows_gpu.items() if k != "rows"}
    log(f"oracle rows ({len(rows)}) in {time.time() - t0:.1f}s: {errs}")
    out["rows_checked"] = len(rows)
    out["rows_rel_fro"] = errs
    return out


INT8_PEAK_TOPS = 5000.0  # MI355X dense int8 MFMA peak


def split_leg(args, h, uvt, dA, dB, dC, flops, plain_ms, nm, rows):
    """Split-precision TRMM with nm moduli on the same inputs: time, the
    call's own error statement (bound, estimate), its error against the FP64
    result, and its sampled rows (checked against exact rows later)."""
    n, m = args.n, args.m
    ws = h.split_workspace(n, m, nm)
    log(f"split: {nm} moduli, workspace {ws / 2**30:.2f} GiB")
    dW, dS = h.alloc(ws), h.alloc(n * m * 8)
    for _ in range(max(1, args.warmup)):
        h.trmm_split(n, m, nm, dA, n, dB, m, dS, m, dW)
    h.sync()
    evs = [(h.event(), h.event()) for _ in range(args.steps)]
    for e0, e1 in evs:
        h.record(e0)
        h.trmm_split(n, m, nm, dA, n, dB, m, dS, m, dW)
        h.record(e1)
    h.sync()
    ms = float(np.mean([h.elapsed_ms(e0, e1) for e0, e1 in evs]))
    bound, estimate = h.split_error(n, m, nm, dW)
    S = dS.download((n, m))
    P = dC.download((n, m))
    err = float(np.linalg.norm(S - P) / np.linalg.norm(P))
    srows = S[rows].copy()
    del S, P
    dW.free()
    dS.free()
    kb = (np.arange(n, dtype=np.float64) // 256) * 256
    int8_ops = 2.0 * m * kb.sum() * nm
    log(f"split: {ms:.2f} ms/step, {flops / (ms * 1e-3) * 1e-12:.2f} TFLOP/s-equivalent, "
        f"rel_fro vs FP64 {err:.3e}, stated bound {bound:.3e} estimate {estimate:.3e}")
    return {"value": round(flops / (ms * 1e-3) * 1e-12, 4), "unit": "TFLOP/s (FP64-equivalent)",
            "ms_per_step": round(ms, 4), "moduli": nm, "workspace_gib": round(ws / 2**30, 3),
            "speedup_vs_plain": round(plain_ms / ms, 4), "rel_fro_vs_fp64": err,
            "stated_error_bound": bound, "stated_error_estimate": estimate,
            "int8_ops_per_step": int8_ops,
            "int8_tops_whole_call": round(int8_ops / (ms * 1e-3) * 1e-12, 2),
            "int8_peak_tops": INT8_PEAK_TOPS}, srows


def batch_leg(args, h, uvt, route):
    """512 uneven TRMMs (orders log-uniform in [64, 4096], B columns uniform in
    [16, 1024], half FP64 / half split class) in one hipGraph launch.  route
    auto: a split-class problem runs in FP64 at these orders (faster in a
    batch below n = 8192, and at least as accurate); route int8: the split
    half takes the int8 path wherever it has a box (n > 256).
    int8_path_share_of_flops says how much of the work the int8 path took."""
    probs_nm = workloads.batch_problems(args.seed, args.batch)
    table = os.path.join(ROOT, "uvatrmm_amd", "tune", "f64_tiles.json")
    table = table if os.path.exists(table) else None
    bufs, probs = [], []
    for k, (n, m, prec) in enumerate(probs_nm):
        lda, ldb = n + (n & 1), m + (m & 1)  # even leading dimensions: 16-byte loads
        dA, dB, dC = h.alloc(n * lda * 8), h.alloc(n * ldb * 8), h.alloc(n * m * 8)
        h.fill_seeded(n, m, args.seed + k, args.spread, dA, lda, dB, ldb)
        bufs += [dA, dB, dC]
        r = uvt.ROUTE_INT8 if (route == "int8" and prec == uvt.PREC_SPLIT) else uvt.ROUTE_AUTO
        probs.append(uvt.Problem(dA.ptr, lda, dB.ptr, ldb, dC.ptr, m, n, m, prec, r))
    h.sync()
    t0 = time.perf_counter()
    batch = uvt.Batch(h, probs, slices=args.split_moduli or 13, table=table)
    build_s = time.perf_counter() - t0
    info = batch.info()
    for _ in range(max(1, args.warmup)):
        batch.run()
    h.sync()
    evs = [(h.event(), h.event()) for _ in range(args.steps)]
    for e0, e1 in evs:
        h.record(e0)
        batch.run()
        h.record(e1)
    h.sync()
    ms_step = float(np.mean([h.elapsed_ms(e0, e1) for e0, e1 in evs]))
    errs = [b for b, _ in batch.errors() if b >= 0]
    ests = [e for _, e in batch.errors() if e >= 0]
    batch.close()
    for b in bufs:
        b.free()
    f64_flops = sum(float(m) * n * n for n, m, p in probs_nm if p == uvt.PREC_F64)
    log(f"batch ({route}): {args.batch} problems, {info['graph_nodes']} graph nodes, "
        f"{ms_step:.2f} ms per launch, {info['flops'] / (ms_step * 1e-3) * 1e-12:.2f} "
        f"TFLOP/s-equivalent, int8 share {info['int8_flops'] / info['flops']:.3f}")
    out = {"problems": args.batch, "route": route,
           "value": round(info["flops"] / (ms_step * 1e-3) * 1e-12, 4),
           "unit": "TFLOP/s (sum m n^2 / launch time)", "ms_per_launch": round(ms_step, 4),
           "flops": info["flops"], "fp64_class_share_of_flops": round(f64_flops / info["flops"], 4),
           "int8_path_share_of_flops": round(info["int8_flops"] / info["flops"], 4),
           "fp64_tiles": info["fp64_tiles"], "graph_nodes": info["graph_nodes"],
           "graph_build_s": round(build_s, 3), "tuning_table": bool(table)}
    if errs:
        out["int8_problems"] = len(errs)
        out["max_stated_error_bound"] = max(errs)
        out["max_stated_error_estimate"] = max(ests)
    return out


def serve_leg(args, h, uvt, big_wait_ms=0):
    """Served stream: TRMM requests with orders log-uniform in [128, 16384], B
    columns uniform in {64..4096}, half FP64 and half split class, inputs
    generated on the device per request, under a 200 GB HBM admission budget;
    requests of m n^2 >= 5e10 (about 0.75 ms of the GPU) run one at a time.
    Capacity is measured first (a burst of requests submitted at once); the
    stream then arrives as a Poisson process at `serve_load` x capacity, and
    latency is arrival (submit) -> result ready in HBM, over the requests
    after the warm-up ones.  big_wait_ms > 0: large requests held back by the
    cap go shortest first, oldest first once they have waited that long."""
    rng = np.random.default_rng(args.seed + 29)
    table = os.path.join(ROOT, "uvatrmm_amd", "tune", "f64_tiles.json")
    table = table if os.path.exists(table) else None

    def draw(count):
        return workloads.serve_sizes(rng, count)

    budget = int(args.serve_budget_gb * 1e9)
    srv = uvt.Server(h, budget, max_slots=args.serve_slots, slices=args.split_moduli or 13,
                     table=table, big_flops=args.serve_big_flops, max_big=args.serve_max_big,
                     big_wait_ms=big_wait_ms)
    seed = [args.seed * 1000003]

    def submit(req, tag):
        n, m, prec = req
        seed[0] += 1
        return srv.submit_seeded(n, m, prec, seed[0], spread=args.spread, tag=tag)

    try:
        burst = draw(args.serve_capacity_reqs)
        for k, r in enumerate(burst):
            submit(r, k)
        srv.drain()
        done = srv.poll(len(burst) + 16)
        span = max(d["t_done"] for d in done) - min(d["t_submit"] for d in done)
        capacity = len(burst) / span
        burst_tflops = sum(float(m) * n * n for n, m, _ in burst) / span * 1e-12
        rate = args.serve_load * capacity
        total = args.serve_warmup + args.serve_requests
        stream = draw(total)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, total))
        log(f"serve: capacity {capacity:.1f} req/s ({burst_tflops:.2f} TFLOP/s burst); "
            f"Poisson stream of {total} at {rate:.1f} req/s")
        t0 = time.perf_counter()
        for k, r in enumerate(stream):
            while True:
                wait = arrivals[k] - (time.perf_counter() - t0)
                if wait <= 0:
                    break
                if wait > 2e-3:
                    time.sleep(wait - 1e-3)
            submit(r, 1_000_000 + k)
            if k % 256 == 0:
                log(f"serve: submitted {k}/{total}")
        srv.drain()
        done = srv.poll(total + 16)
        st = srv.stats()
    finally:
        srv.close()
    assert len(done) == total and all(d["status"] == uvt.OK for d in done)
    meas = [d for d in done if d["tag"] - 1_000_000 >= args.serve_warmup]
    lat = np.sort([d["t_done"] - d["t_submit"] for d in meas])

    def pct(v, q):  # nearest-rank percentile of a sorted array
        return float(v[max(1, int(np.ceil(q * len(v)))) - 1])

    rank = lambda q: pct(lat, q)
    # latency by size class (m n^2): small < 1e9 <= medium < big_flops <= large
    classes = {}
    for name, lo, hi in (("small", 0.0, 1e9), ("medium", 1e9, args.serve_big_flops),
                         ("large", args.serve_big_flops, float("inf"))):
        sel = [d for d in meas
               if lo <= float(stream[d["tag"] - 1_000_000][1]) * stream[d["tag"] - 1_000_000][0] ** 2 < hi]
        if sel:
            v = np.sort([d["t_done"] - d["t_submit"] for d in sel])
            classes[name] = {"requests": len(sel), "p50_ms": round(pct(v, 0.5) * 1e3, 4),
                             "p99_ms": round(pct(v, 0.99) * 1e3, 4),
                             "max_ms": round(float(v[-1]) * 1e3, 4)}
    int8 = [d for d in meas if d["path"] == 1]
    span = max(d["t_done"] for d in meas) - min(d["t_submit"] for d in meas)
    fl = sum(float(stream[d["tag"] - 1_000_000][1]) * stream[d["tag"] - 1_000_000][0] ** 2
             for d in meas)
    out = {"requests": len(meas), "warmup_requests": args.serve_warmup,
           "p50_ms": round(rank(0.50) * 1e3, 4), "p99_ms": round(rank(0.99) * 1e3, 4),
           "p999_ms": round(rank(0.999) * 1e3, 4),
           "mean_ms": round(float(lat.mean()) * 1e3, 4), "max_ms": round(float(lat[-1]) * 1e3, 4),
           "throughput_tflops": round(fl / span * 1e-12, 4),
           "requests_per_s": round(len(meas) / span, 2),
           "arrival_rate_per_s": round(rate, 2), "load": args.serve_load,
           "capacity_per_s": round(capacity, 2), "capacity_tflops": round(burst_tflops, 4),
           "slots": args.serve_slots, "budget_gb": args.serve_budget_gb,
           "large_request_flops": args.serve_big_flops, "max_large_in_flight": args.serve_max_big,
           "large_order": ("first come first served" if big_wait_ms == 0 else
                           f"shortest first; oldest first once waited {big_wait_ms} ms"),
           "peak_arena_gb": round(st["peak_bytes"] / 1e9, 3),
           "peak_inflight": st["peak_inflight"],
           "split_class_share": round(float(np.mean([p for _, _, p in stream])), 3),
           "int8_path_share": round(len(int8) / len(meas), 4),
           "by_size_class": classes,
           "int8_max_stated_error_bound": max((d["err_bound"] for d in int8), default=None),
           "int8_max_stated_error_estimate": max((d["err_estimate"] for d in int8), default=None),
           "latency": "submit -> result ready in HBM (inputs generated on device per request)"}
    log(f"serve: p50 {out['p50_ms']} ms p99 {out['p99_ms']} ms, {out['throughput_tflops']} "
        f"TFLOP/s, {out['requests_per_s']} req/s, peak {st['peak_inflight']} in flight")
    return out


def vendor_leg(n=16384):
    """Sustained rates of the vendor GEMMs on this GPU through torch
    (hipBLASLt / rocBLAS), as practical ceilings beside the published peaks:
    int8 x int8 -> int32 (torch._int_mm, B column-major, the faster of its two
    layouts) and an FP64 matmul, square n = 16384, HIP events on torch's
    stream."""
    import torch

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    dev = "cuda:0"
    a = torch.randint(-128, 128, (n, n), dtype=torch.int8, device=dev)
    b = torch.randint(-128, 128, (n, n), dtype=torch.int8, device=dev).t().contiguous().t()
    ms_i8 = timed(lambda: torch._int_mm(a, b), 10)
    del a, b
    a = torch.randn(n, n, dtype=torch.float64, device=dev)
    b = torch.randn(n, n, dtype=torch.float64, device=dev)
    ms_f64 = timed(lambda: a @ b, 3)
    del a, b
    torch.cuda.empty_cache()
    ops = 2.0 * n ** 3
    out = {"n": n, "int8_ms": round(ms_i8, 4), "int8_tops": round(ops / (ms_i8 * 1e-3) * 1e-12, 1),
           "fp64_ms": round(ms_f64, 4), "fp64_tflops": round(ops / (ms_f64 * 1e-3) * 1e-12, 2),
           "via": "torch._int_mm and torch.matmul (hipBLASLt / rocBLAS)"}
    log(f"vendor GEMMs at n={n}: int8 {out['int8_tops']} TOP/s, FP64 {out['fp64_tflops']} TFLOP/s")
    return out


def configs_leg(args, h, uvt):
    """The reference's other configurations, FP64: N=2048 (M=N, reg, boxes,
    P=4) and N=16384 (M=N, balanced, trapezoid, P=16), seeded inputs."""
    out = []
    for n, p, mp, shape in ((2048, 4, "reg", "boxes"), (16384, 16, "balanced", "trapezoid")):
        m = n
        dA, dB, dC = h.alloc(n * n * 8), h.alloc(n * m * 8), h.alloc(n * m * 8)
        h.fill_seeded(n, m, args.seed, args.spread, dA, n, dB, m)
        plan = uvt.Plan(n, m, p, mp, shape)
        for _ in range(max(1, args.warmup)):
            h.trmm_f64(plan, dA, n, dB, m, dC, m)
        h.sync()
        evs = [(h.event(), h.event()) for _ in range(args.steps)]
        for e0, e1 in evs:
            h.record(e0)
            h.trmm_f64(plan, dA, n, dB, m, dC, m)
            h.record(e1)
        h.sync()
        ms = float(np.mean([h.elapsed_ms(e0, e1) for e0, e1 in evs]))
        tf = plan.flops / (ms * 1e-3) * 1e-12
        log(f"config N={n} {mp} {shape} P={p}: {ms:.3f} ms, {tf:.2f} TFLOP/s")
        out.append({"workload": f"trmm_f64 N=M={n} {mp} {shape} P={p}", "ms": round(ms, 4),
                    "tflops": round(tf, 4), "frac": round(tf / FP64_PEAK_TFLOPS, 4),
                    "phases": plan.phases})
        plan.close()
        for b in (dA, dB, dC):
            b.free()
    return out


def sample_rows(n, bands, mapping):
    """Rows of C checked against exact rows: the ends of the first, middle and
    last 256-row tiles, the 8192-row boundaries, every band start, and a few
    seeded others."""
    import uvatrmm_amd as uvt

    rows = {0, 255, 256, n // 4 - 1, n // 4, n // 2 - 1, n // 2, n - 257, n - 256, n - 1}
    rows |= {b for b, _ in uvt.plan_bands(n, bands, mapping)}
    rng = np.random.default_rng(n + 5)
    rows |= set(int(x) for x in rng.integers(0, n, 8))
    return sorted(r for r in rows if 0 <= r < n)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--m", type=int, default=8192)
    ap.add_argument("--bands", type=int, default=16)
    ap.add_argument("--map", default="balanced", choices=["reg", "balanced"])
    ap.add_argument("--shape", default="trapezoid", choices=["boxes", "trapezoid"])
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--spread", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-cols-per-rank", type=int, default=128)
    ap.add_argument("--batch", type=int, default=512, help="problems of the batch leg (0 = skip)")
    ap.add_argument("--split-moduli", type=int, default=12,
                    help="moduli of the split-precision leg (0 = skip it); 12 keeps the "
                         "error below 1e-12 at N=32768")
    ap.add_argument("--split-moduli-alt", type=int, nargs="*", default=[13],
                    help="further moduli counts measured beside it (time and error)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the N=2048 and N=16384 reference configurations")
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="measured requests of the served stream (0 = skip)")
    ap.add_argument("--serve-warmup", type=int, default=100)
    ap.add_argument("--serve-capacity-reqs", type=int, default=300)
    ap.add_argument("--serve-load", type=float, default=0.7)
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-big-flops", type=float, default=5e10,
                    help="m n^2 from which a served request counts as large")
    ap.add_argument("--serve-max-big", type=int, default=1,
                    help="large requests in flight at once (0 = no cap)")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-alt-wait-ms", type=int, default=150,
                    help="second serve leg on the same stream with the large requests shortest "
                         "first, except those waiting this many ms or longer (0 = skip it)")
    ap.add_argument("--no-vendor-ceiling", dest="vendor_ceiling", action="store_false",
                    help="skip timing the vendor int8 and FP64 GEMMs (torch) at n=16384")
    ap.add_argument("--vendor-only", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.vendor_only:
        print(json.dumps(vendor_leg()), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the TRMM path shards by columns of B with no "
              "reduction, and no multi-GPU run is part of this build.", file=sys.stderr)
        sys.exit(2)

    import uvatrmm_amd as uvt

    n, m = args.n, args.m
    h = uvt.Handle(0)
    log(f"device cus={h.cus}; allocating N={n} M={m}")
    dA, dB, dC = h.alloc(n * n * 8), h.alloc(n * m * 8), h.alloc(n * m * 8)
    h.fill_seeded(n, m, args.seed, args.spread, dA, n, dB, m)
    plan = uvt.Plan(n, m, args.bands, args.map, args.shape)
    flops = plan.flops
    h.sync()
    log(f"plan: {plan.phases} launch(es), tiles={[len(plan.tiles(p)) for p in range(plan.phases)]}, "
        f"issued/algorithmic={plan.issued_flops / flops:.4f}")
    for _ in range(args.warmup):
        h.trmm_f64(plan, dA, n, dB, m, dC, m)
    h.sync()
    evs = [(h.event(), h.event()) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for e0, e1 in evs:
        h.record(e0)
        h.trmm_f64(plan, dA, n, dB, m, dC, m)
        h.record(e1)
    h.sync()
    wall = time.perf_counter() - t0
    step_ms = [h.elapsed_ms(e0, e1) for e0, e1 in evs]
    mean_ms = float(np.mean(step_ms))
    value = flops / (wall / args.steps) * 1e-12
    achieved = flops / (mean_ms * 1e-3) * 1e-12
    log(f"steps ms: min {min(step_ms):.2f} mean {mean_ms:.2f} max {max(step_ms):.2f}; "
        f"wall/step {wall / args.steps * 1e3:.2f} ms; {achieved:.2f} TFLOP/s")
    norm = h.inf_norm(n, m, dC, m)
    rows = sample_rows(n, args.bands, args.map)
    rows_gpu = {"rows": rows}
    full = dC.download((n, m)) if m * n * 8 <= (1 << 34) else None
    if full is not None:
        rows_gpu["fp64"] = full[rows].copy()
        del full

    split, split_alt = None, []
    if args.split_moduli > 0:
        split, rows_gpu[f"split{args.split_moduli}"] = split_leg(args, h, uvt, dA, dB, dC, flops,
                                                                  mean_ms, args.split_moduli, rows)
        for nm in args.split_moduli_alt:
            if nm != args.split_moduli:
                leg, rows_gpu[f"split{nm}"] = split_leg(args, h, uvt, dA, dB, dC, flops, mean_ms,
                                                        nm, rows)
                split_alt.append(leg)

    batch = batch_leg(args, h, uvt, "auto") if args.batch > 0 else None
    batch_forced = batch_leg(args, h, uvt, "int8") if args.batch > 0 else None

    other = configs_leg(args, h, uvt) if not args.no_configs else None

    serve = serve_leg(args, h, uvt) if args.serve_requests > 0 else None
    # the same stream with the large requests shortest first (a lower p99 and
    # mean for a longer wait of the largest requests: p999 and max reported)
    serve_sf = (serve_leg(args, h, uvt, args.serve_alt_wait_ms)
                if args.serve_requests > 0 and args.serve_alt_wait_ms > 0 else None)
    vendor = None
    if args.vendor_ceiling:
        # a child process: torch brings its own HIP runtime, which does not
        # share a process with the one this library was built against
        import subprocess
        log("vendor GEMMs: timing in a child process")
        try:
            r = subprocess.run([sys.executable, os.path.abspath(__file__), "--vendor-only"],
                               capture_output=True, text=True, timeout=300)
            if r.returncode == 0 and r.stdout.strip():
                vendor = json.loads(r.stdout.strip().splitlines()[-1])
            else:
                log("vendor GEMM timing failed:", r.stderr.strip()[-400:])
        except Exception as e:  # a reported ceiling, never the measured path
            log("vendor GEMM timing failed:", repr(e))

    cpu = None
    if not args.no_cpu:
        cols = min(m, 16 * args.cpu_cols_per_rank)
        Cg = None
        full = dC.download((n, m)) if m * n * 8 <= (1 << 34) else None
        if full is not None:
            Cg = np.ascontiguousarray(full[:, :cols])
            del full
        try:
            cpu = cpu_baseline(args, Cg, rows_gpu)
            log(f"cpu baseline {cpu['value']:.4f} TFLOP/s on {cpu['cores']} cores; "
                f"gpu vs cpu rel_fro {cpu['gpu_vs_cpu_rel_fro']}")
            if split is not None:
                split["rel_fro_vs_exact_rows"] = cpu["rows_rel_fro"].get(f"split{args.split_moduli}")
            for leg in split_alt:
                leg["rel_fro_vs_exact_rows"] = cpu["rows_rel_fro"].get(f"split{leg['moduli']}")
            if other is not None:
                for g, c in zip(other, cpu["other_configs"]):
                    g["cpu_baseline"] = c
        except Exception as e:  # the baseline is reported, never the measured path
            log("cpu baseline failed:", repr(e))

    # the kernels this plan runs on this GPU (whole tiles, stream-K, or cut reductions)
    wg, parts = h.plan_launch_info(plan)
    cfg0 = int(plan.tiles(0)["cfg"][0]) if len(plan.tiles(0)) else 0
    tile, waves = {0: ("128,128", "8 waves of 32x64"), 1: ("64,64", "4 waves of 32x32"),
                   2: ("128,32", "4 waves of 32x32"), 4: ("256,128", "16 waves of 32x64")}[cfg0]
    vec = "true" if n % 2 == 0 and m % 2 == 0 else "false"
    launches = plan.phases
    if parts and cfg0 == 0 and plan.phases == 1:
        kernel = f"k_trmm_sk_f64<128,128,{vec}> (stream-K, {wg} workgroups) + k_sk_fixup"
        launches = 2
    elif parts:
        kernel = f"k_trmm_tiles_f64<{tile},{vec}> ({waves}; reductions cut into pieces) + k_add_parts"
        launches = plan.phases + 1
    else:
        kernel = f"k_trmm_tiles_f64<{tile},{vec}> ({waves})"
    out = {
        "metric": "TRMM TFLOP/s + % MFMA peak at N=32768 (plain FP64 = value; split in 'split')",
        "value": round(value, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded U[-1,1) x 2^[-8,8], generated on device)",
        "config": {"workload": f"trmm_f64 left lower non-unit N={n} M={m} {args.map} "
                               f"{args.shape} P={args.bands}",
                   "n": n, "m": m, "bands": args.bands, "map": args.map, "shape": args.shape},
        "roofline": {"bound": "mfma", "kernel": kernel, "achieved": round(achieved, 4),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": MEASURED_TRAFFIC.get((n, m, args.bands, args.map, args.shape)),
                     "launches_per_step": launches,
                     "algorithmic_flops_per_step": flops},
        "cpu_baseline": cpu,
        "split": split,
        "split_other_moduli": split_alt,
        "batch": batch,
        "batch_forced_split": batch_forced,
        "other_configs": other,
        "serve": serve,
        "serve_shortest_first": serve_sf,
        "vendor_gemm": vendor,
        "norm": norm,
    }
    print(json.dumps(out), flush=True)
    for b in (dA, dB, dC):
        b.free()
    plan.close()
    h.close()


if __name__ == "__main__":
    main()
