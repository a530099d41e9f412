"""The TRMM server (uvt_serve_*): requests of mixed sizes and classes admitted
by HBM budget and free slots.  A request's result does not depend on what
else is in flight (bit for bit equal to the same request served alone), the
arena never holds more than the budget, oversize requests are refused, and
every result agrees with the FP64 oracle within its class's tolerance."""
import sys
import time

import numpy as np
import pytest

from conftest import rel_fro

pytestmark = pytest.mark.gpu

GiB = 1 << 30


def say(*a):
    print("[gpu-serve]", *a, flush=True)
    sys.stdout.flush()


def mixed_requests(count, seed, nmax=3000, mmax=700):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(48), np.log(nmax))))
        m = int(np.exp(rng.uniform(np.log(8), np.log(mmax))))
        prec = 1 if rng.random() < 0.5 else 0
        out.append((n, m, prec, 5000 + k))
    return out


def serve(uvt, gpu, reqs, budget, slots):
    srv = uvt.Server(gpu, budget, max_slots=slots, slices=13)
    outs = [np.full((n, m), np.nan) for n, m, _, _ in reqs]
    try:
        for k, (n, m, prec, seed) in enumerate(reqs):
            srv.submit_seeded(n, m, prec, seed, spread=1, C=outs[k], tag=k)
            if slots == 1:
                srv.drain()
        srv.drain()
        done = srv.poll()
        st = srv.stats()
    finally:
        srv.close()
    assert len(done) == len(reqs)
    assert all(d["status"] == uvt.OK for d in done), [d for d in done if d["status"]]
    assert sorted(d["tag"] for d in done) == list(range(len(reqs)))
    for d in done:
        assert d["t_submit"] <= d["t_admit"] <= d["t_done"]
    return outs, st, {d["tag"]: d for d in done}


def test_serve_alone_equals_mix(uvt, gpu, oracle):
    # two FP64 requests with few, deep 128x128 tiles (their products run
    # stream-K), then two split-class requests large enough for the int8 path
    sk = [(8165, 939, 0, 6003), (5000, 2000, 0, 6004)]
    reqs = mixed_requests(40, seed=11) + sk + [(4096, 1024, 1, 6001), (6144, 640, 1, 6002)]
    say("requests", len(reqs), "alone")
    alone, st1, _ = serve(uvt, gpu, reqs, 8 * GiB, 1)
    assert st1["peak_inflight"] == 1
    say("mix, 8 slots")
    mix, st8, comp = serve(uvt, gpu, reqs, 8 * GiB, 8)
    say("stats", st8)
    assert st8["peak_inflight"] > 1
    assert st8["completed"] == st8["submitted"] == len(reqs)
    for k, (n, m, prec, seed) in enumerate(reqs):
        assert np.array_equal(alone[k], mix[k]), (k, n, m, prec)
    for k, (n, m, prec, seed) in enumerate(reqs[:12]):
        ref = oracle.trmm(oracle.seeded_A(n, seed, 1), oracle.seeded_B(n, m, seed, 1))
        err = rel_fro(mix[k], ref)
        assert err <= (1e-13 if prec == 0 else 1e-12), (k, n, m, prec, err)
    for k in range(40, 42):  # the stream-K requests: exact (double-double) rows
        n, m, prec, seed = reqs[k]
        rows = np.unique(np.concatenate([[0, 127, 128, n - 1], np.arange(0, n, 400)]))
        A, B = oracle.seeded_A(n, seed, 1), oracle.seeded_B(n, m, seed, 1)
        err = rel_fro(mix[k][rows], oracle.trmm_rows_dd(A, B, rows))
        say("stream-k request", (n, m), "err", err)
        assert err <= 1e-13
    # the two large split-class requests ran on the int8 path and state their
    # error: the bound covers the error against the oracle; every other
    # request ran in FP64 and states none
    for k, (n, m, prec, seed) in enumerate(reqs):
        int8 = k >= len(reqs) - 2
        assert comp[k]["path"] == (1 if int8 else 0), (k, n, m, prec)
        if not int8:
            assert comp[k]["err_bound"] == -1.0
            continue
        # exact rows (double-double sums) of a sample: at 13 moduli the split
        # error is below the rounding error of a plain FP64 product
        rng = np.random.default_rng(k)
        rows = np.unique(np.concatenate([np.arange(0, n, 256), rng.integers(0, n, 160)]))
        A, B = oracle.seeded_A(n, seed, 1), oracle.seeded_B(n, m, seed, 1)
        err = rel_fro(mix[k][rows], oracle.trmm_rows_dd(A, B, rows))
        say("int8 request", (n, m), "err", err, "bound", comp[k]["err_bound"],
            "estimate", comp[k]["err_estimate"])
        assert err <= 1e-12
        assert err <= comp[k]["err_bound"]
        assert 0.25 <= err / comp[k]["err_estimate"] <= 4.0


def test_serve_budget_admission(uvt, gpu):
    reqs = mixed_requests(30, seed=12, nmax=2500, mmax=600)
    need = max(uvt.serve_request_bytes(n, m, p) for n, m, p, _ in reqs)
    budget = int(need * 1.5)
    say("budget", budget, "largest", need)
    outs, st, _ = serve(uvt, gpu, reqs, budget, 16)
    assert st["peak_bytes"] <= budget
    alone, _, _ = serve(uvt, gpu, reqs[:6], 8 * GiB, 1)
    for k in range(6):
        assert np.array_equal(alone[k], outs[k])


def test_serve_rejects_oversize(uvt, gpu):
    srv = uvt.Server(gpu, 64 << 20, max_slots=2)
    try:
        with pytest.raises(uvt.UvtError) as e:
            srv.submit_seeded(4096, 1024, uvt.PREC_F64, 1)
        assert e.value.status == uvt.ERR_BUDGET
        C = np.zeros((256, 100))
        srv.submit_seeded(256, 100, uvt.PREC_F64, 1, C=C)
        srv.drain()
        st = srv.stats()
        assert st["rejected"] == 1 and st["completed"] == 1
        assert st["p50_s"] <= st["p99_s"] <= st["max_s"]
        assert np.isfinite(C).all()
    finally:
        srv.close()


def test_serve_host_arrays(uvt, gpu, oracle):
    srv = uvt.Server(gpu, 2 * GiB, max_slots=4)
    cases = [(300, 77, 0), (1030, 130, 1), (900, 64, 1), (65, 9, 0)]
    outs = []
    try:
        for k, (n, m, prec) in enumerate(cases):
            A = oracle.seeded_A(n, 70 + k, 1)
            B = oracle.seeded_B(n, m, 70 + k, 1)
            C = np.zeros((n, m))
            srv.submit(A, B, prec, C=C, tag=k)
            outs.append((A, B, C, prec))
        srv.drain()
        assert len(srv.poll()) == len(cases)
    finally:
        srv.close()
    for A, B, C, prec in outs:
        err = rel_fro(C, oracle.trmm(A, B))
        assert err <= (1e-13 if prec == 0 else 1e-12), (A.shape, prec, err)


def test_serve_device_operands(uvt, gpu, oracle):
    """Operands already resident in HBM are read and written in place and the
    arena is charged only for the split workspace."""
    cases = [(1100, 96, 1), (500, 40, 0), (2048, 256, 0), (4096, 1024, 1)]  # last: int8 path
    srv = uvt.Server(gpu, 1 << 30, max_slots=4)
    bufs = []
    try:
        for k, (n, m, prec) in enumerate(cases):
            A = oracle.seeded_A(n, 90 + k, 1)
            B = oracle.seeded_B(n, m, 90 + k, 1)
            dA, dB, dC = gpu.array(A), gpu.array(B), gpu.alloc(n * m * 8)
            srv.submit_device(n, m, prec, dA, dB, dC, tag=k)
            bufs.append((A, B, dC, prec))
        srv.drain()
        st = srv.stats()
    finally:
        srv.close()
    ws = max(uvt.serve_request_bytes(n, m, p) - uvt.serve_request_bytes(n, m, uvt.PREC_F64)
             for n, m, p in cases)
    assert 0 < st["peak_bytes"] <= len(cases) * ws
    for A, B, dC, prec in bufs:
        n, m = B.shape
        err = rel_fro(dC.download((n, m)), oracle.trmm(A, B))
        assert err <= (1e-13 if prec == 0 else 1e-12), (n, m, prec, err)


def test_serve_large_request_cap(uvt, gpu, oracle):
    """With max_big = 1, requests of m n^2 >= big_flops run one at a time
    (their admit..done intervals do not overlap), a smaller request queued
    behind a held-back large one is admitted before it, and every result is
    bit for bit the same request served alone."""
    # The first one keeps the GPU busy for about 4 ms: longer than the
    # submissions and than the scheduler's own launch of it (about 1.5 ms
    # has been seen for the first launch of a kernel in a process), so the
    # cap is still holding request 1 back when the small ones are looked at.
    reqs = [(8192, 4096, 0, 7001), (1400, 700, 1, 7002), (300, 50, 0, 7003), (1600, 500, 0, 7004),
            (200, 40, 1, 7005), (1300, 800, 1, 7006), (250, 30, 0, 7007)]
    big = 5e8  # the four large ones have m n^2 >= 1.28e9, the small ones <= 4.5e6
    srv = uvt.Server(gpu, 8 * GiB, max_slots=8, slices=13, big_flops=big, max_big=1)
    outs = [np.full((n, m), np.nan) for n, m, _, _ in reqs]
    try:
        # all submitted before the scheduler can retire the first large one
        for k, (n, m, prec, seed) in enumerate(reqs):
            srv.submit_seeded(n, m, prec, seed, spread=1, C=outs[k], tag=k)
        srv.drain()
        done = {d["tag"]: d for d in srv.poll()}
    finally:
        srv.close()
    assert sorted(done) == list(range(len(reqs))) and all(d["status"] == uvt.OK for d in done.values())
    times = {k: (round(d["t_submit"] * 1e3, 3), round(d["t_admit"] * 1e3, 3), round(d["t_done"] * 1e3, 3))
             for k, d in done.items()}  # ms: submit, admit, done
    print("[gpu] serve cap times", times, flush=True)
    large = sorted((done[k] for k, (n, m, _, _) in enumerate(reqs) if m * n * n >= big),
                   key=lambda d: d["t_admit"])
    assert len(large) == 4
    assert [d["tag"] for d in large] == [0, 1, 3, 5], times  # large ones keep their order
    for a, b in zip(large, large[1:]):
        assert b["t_admit"] >= a["t_done"], times
    assert done[2]["t_submit"] < done[0]["t_done"], times  # the scenario: 0 still running
    assert done[2]["t_admit"] < done[1]["t_admit"], times  # the small one passed the held-back large one
    alone, _, _ = serve(uvt, gpu, reqs, 8 * GiB, 1)
    for k, (n, m, prec, seed) in enumerate(reqs):
        assert np.array_equal(alone[k], outs[k]), (k, n, m)
    for k in (2, 6):
        n, m, prec, seed = reqs[k]
        ref = oracle.trmm(oracle.seeded_A(n, seed, 1), oracle.seeded_B(n, m, seed, 1))
        assert rel_fro(outs[k], ref) <= 1e-13


@pytest.mark.parametrize("wait_ms,expect", [
    (0, "arrival"),       # first come first served
    (10000, "size"),      # shortest first; none waits 10 s here
    (1, "arrival"),       # shortest first, but all have waited 1 ms: oldest first
])
def test_serve_large_order(uvt, gpu, wait_ms, expect):
    """big_wait_ms: which large request held back by the cap goes next.  A
    large request keeps the GPU busy for about 4 ms while four more, their
    m n^2 3x apart, arrive largest first; they are then admitted in arrival
    order or by size as big_wait_ms says.  Results bit for bit as served
    alone."""
    reqs = [(8192, 4096, 0, 7101), (3000, 3000, 0, 7102), (2000, 2250, 1, 7103), (1440, 1450, 0, 7104),
            (1000, 1000, 1, 7105), (300, 50, 0, 7106)]
    big = 5e8  # large: m n^2 = 2.7e11, then 2.7e10, 9.0e9, 3.0e9, 1.0e9; the last one is small
    srv = uvt.Server(gpu, 8 * GiB, max_slots=8, slices=13, big_flops=big, max_big=1,
                     big_wait_ms=wait_ms)
    outs = [np.full((n, m), np.nan) for n, m, _, _ in reqs]
    try:
        # the first large request alone (the scheduler admits it at once),
        # then the others while it runs
        srv.submit_seeded(*reqs[0][:3], reqs[0][3], spread=1, C=outs[0], tag=0)
        time.sleep(0.001)
        for k, (n, m, prec, seed) in enumerate(reqs[1:], 1):
            srv.submit_seeded(n, m, prec, seed, spread=1, C=outs[k], tag=k)
        srv.drain()
        done = {d["tag"]: d for d in srv.poll()}
    finally:
        srv.close()
    assert sorted(done) == list(range(len(reqs))) and all(d["status"] == uvt.OK for d in done.values())
    flops = {k: m * n * n for k, (n, m, _, _) in enumerate(reqs)}
    large = [d["tag"] for d in sorted((done[k] for k in flops if flops[k] >= big), key=lambda d: d["t_admit"])]
    times = {k: (round(d["t_submit"] * 1e3, 3), round(d["t_admit"] * 1e3, 3), round(d["t_done"] * 1e3, 3))
             for k, d in done.items()}
    say("large admitted in order", large, "wait_ms", wait_ms, "times", times)
    # the scenario: the others were all queued while the first one ran
    assert large[0] == 0 and all(done[k]["t_submit"] < done[0]["t_done"] for k in large), times
    want = [1, 2, 3, 4] if expect == "arrival" else [4, 3, 2, 1]
    assert large[1:] == want, (large, times)
    alone, _, _ = serve(uvt, gpu, reqs, 8 * GiB, 1)
    for k, (n, m, prec, seed) in enumerate(reqs):
        assert np.array_equal(alone[k], outs[k]), (k, n, m)


def test_serve_config_checks(uvt, gpu):
    for kw in ({"max_big": -1}, {"big_flops": -1.0}, {"big_wait_ms": -1}):
        with pytest.raises(uvt.UvtError) as e:
            uvt.Server(gpu, 1 << 30, max_slots=2, **kw)
        assert e.value.status == uvt.ERR_ARG
