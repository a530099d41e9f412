"""Batched TRMMs in one hipGraph launch (persistent FP64 queue + batched split
stages): every result equals the same problem run alone (bit for bit: the
order of every sum is fixed per problem) and the FP64 oracle within the
path's tolerance; graph replays are identical."""
import sys

import numpy as np
import pytest

from conftest import rel_fro

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-batch]", *a, flush=True)
    sys.stdout.flush()


def make_problems(uvt, gpu, oracle, sizes, seed=0):
    rng = np.random.default_rng(seed)
    probs, host, bufs = [], [], []
    for k, (n, m) in enumerate(sizes):
        prec = uvt.PREC_SPLIT if k % 2 else uvt.PREC_F64
        A = oracle.seeded_A(n, 1000 + k, 1)
        B = oracle.seeded_B(n, m, 1000 + k, 1)
        dA, dB, dC = gpu.array(A), gpu.array(B), gpu.alloc(n * m * 8)
        bufs.append((dA, dB, dC))
        host.append((A, B, prec))
        probs.append(uvt.Problem(dA.ptr, n, dB.ptr, m, dC.ptr, m, n, m, prec, 0))
    return probs, host, bufs


# Odd positions are split class.  Below n = 8192 a split-class problem runs in
# FP64 in a batch; (8200, 260) and (8448, 300) take the int8 path.
SIZES = [(64, 16), (100, 33), (257, 64), (300, 1000), (513, 128), (700, 250), (1024, 512),
         (1500, 48), (2049, 300), (4096, 96), (80, 700), (1200, 1024), (130, 17), (600, 600),
         (70, 30), (8200, 260), (90, 20), (8448, 300)]


def test_batch_matches_alone_and_oracle(uvt, gpu, oracle):
    probs, host, bufs = make_problems(uvt, gpu, oracle, SIZES)
    say("batch of", len(probs))
    batch = uvt.Batch(gpu, probs, slices=13)
    info = batch.info()
    say("info", info)
    assert info["flops"] == sum(float(m) * n * n for n, m in SIZES)
    assert info["int8_flops"] == 260.0 * 8200 * 8200 + 300.0 * 8448 * 8448
    batch.run()
    gpu.sync()
    results = [bufs[k][2].download((n, m)) for k, (n, m) in enumerate(SIZES)]
    for k, (n, m) in enumerate(SIZES):
        A, B, prec = host[k]
        ref = oracle.trmm(A, B)
        err = rel_fro(results[k], ref)
        tol = 1e-13 if prec == uvt.PREC_F64 else 1e-12
        assert err <= tol, (k, n, m, prec, err)
        # alone: the same problem as a batch of one
        dA, dB, _ = bufs[k]
        dC = gpu.alloc(n * m * 8)
        one = uvt.Batch(gpu, [uvt.Problem(dA.ptr, n, dB.ptr, m, dC.ptr, m, n, m, prec, 0)])
        one.run()
        gpu.sync()
        one.close()
        np.testing.assert_array_equal(dC.download((n, m)), results[k], err_msg=f"problem {k}")
        # FP64 problems also equal the single-problem FP64 entry point
        if prec == uvt.PREC_F64:
            plan = uvt.Plan(n, m, 1, "reg", "trapezoid")
            gpu.trmm_f64(plan, dA, n, dB, m, dC, m)
            gpu.sync()
            plan.close()
            np.testing.assert_array_equal(dC.download((n, m)), results[k], err_msg=f"f64 {k}")
        dC.free()
    # replays, with host work and synchronisation in between, give the same bits
    for rep in range(3):
        for _, _, dC in bufs:
            dC.zero()
        batch.run()
        gpu.sync()
        for k, (n, m) in enumerate(SIZES):
            np.testing.assert_array_equal(bufs[k][2].download((n, m)), results[k],
                                          err_msg=f"replay {rep} problem {k}")
    batch.close()
    for t in bufs:
        for b in t:
            b.free()


def test_batch_update_new_pointers(uvt, gpu, oracle):
    sizes = [(300, 40), (600, 200), (1000, 100), (513, 513)]
    probs, host, bufs = make_problems(uvt, gpu, oracle, sizes, seed=1)
    batch = uvt.Batch(gpu, probs)
    batch.run()
    gpu.sync()
    # new B and C buffers with other data
    new, keep = [], []
    for k, (n, m) in enumerate(sizes):
        B2 = oracle.seeded_B(n, m, 77 + k, 0)
        dB2, dC2 = gpu.array(B2), gpu.alloc(n * m * 8)
        keep.append((dB2, dC2, B2))
        new.append(uvt.Problem(bufs[k][0].ptr, n, dB2.ptr, m, dC2.ptr, m, n, m, host[k][2], 0))
    batch.update(new)
    batch.run()
    gpu.sync()
    for k, (n, m) in enumerate(sizes):
        ref = oracle.trmm(host[k][0], keep[k][2])
        tol = 1e-13 if host[k][2] == uvt.PREC_F64 else 1e-12
        assert rel_fro(keep[k][1].download((n, m)), ref) <= tol
    with pytest.raises(uvt.UvtError):
        batch.update(new[:2])
    batch.close()
    for t in bufs:
        for b in t:
            b.free()
    for dB2, dC2, _ in keep:
        dB2.free()
        dC2.free()


def test_batch_update_one_int8_problem(uvt, gpu, oracle):
    """A batch whose only int8-path problem is (8200, 260) beside a small FP64
    one: after uvt_batch_update with new B and C buffers, the captured graph's
    split stages must read the new B and write the new C (the int8 problem's
    stages read it from the device descriptor table), checked against the
    oracle; the old C must stay untouched by the second run."""
    sizes = [(300, 40), (8200, 260)]
    precs = [uvt.PREC_F64, uvt.PREC_SPLIT]
    host, bufs, probs = [], [], []
    for k, ((n, m), prec) in enumerate(zip(sizes, precs)):
        A = oracle.seeded_A(n, 300 + k, 1)
        B = oracle.seeded_B(n, m, 300 + k, 1)
        dA, dB, dC = gpu.array(A), gpu.array(B), gpu.alloc(n * m * 8)
        host.append((A, B))
        bufs.append((dA, dB, dC))
        probs.append(uvt.Problem(dA.ptr, n, dB.ptr, m, dC.ptr, m, n, m, prec, 0))
    batch = uvt.Batch(gpu, probs, slices=13)
    assert batch.info()["int8_flops"] == 260.0 * 8200 * 8200
    batch.run()
    gpu.sync()
    first = [bufs[k][2].download((n, m)) for k, (n, m) in enumerate(sizes)]
    new, keep = [], []
    for k, ((n, m), prec) in enumerate(zip(sizes, precs)):
        B2 = oracle.seeded_B(n, m, 900 + k, 1)
        dB2, dC2 = gpu.array(B2), gpu.alloc(n * m * 8)
        keep.append((dB2, dC2, B2))
        new.append(uvt.Problem(bufs[k][0].ptr, n, dB2.ptr, m, dC2.ptr, m, n, m, prec, 0))
    batch.update(new)
    batch.run()
    gpu.sync()
    errs = batch.errors()
    for k, (n, m) in enumerate(sizes):
        ref = oracle.trmm(host[k][0], keep[k][2])
        err = rel_fro(keep[k][1].download((n, m)), ref)
        say("updated problem", k, (n, m), "err", err, "stated", errs[k])
        assert err <= (1e-13 if precs[k] == uvt.PREC_F64 else 1e-12)
        np.testing.assert_array_equal(bufs[k][2].download((n, m)), first[k])
    assert errs[0] == (-1.0, -1.0)
    assert errs[1][0] >= rel_fro(keep[1][1].download(sizes[1]), oracle.trmm(host[1][0], keep[1][2]))
    batch.close()
    for t in bufs:
        for b in t:
            b.free()
    for dB2, dC2, _ in keep:
        dB2.free()
        dC2.free()


def test_bench_batch_both_routes(uvt, gpu, oracle):
    """The benchmark's 512-problem batch (seed 3), with the shipped tuning
    table, routed by measured speed (every problem in FP64 at these orders)
    and with its split half forced onto the int8 path: FP64-class results
    are the same bits in both; each forced split result is within 1e-12 of
    the routed FP64 one; a sample of 24 problems (the 12 largest, 12 seeded
    others) is checked against the oracle, and every int8-path problem's
    stated bound covers its error against the oracle on that sample."""
    from uvatrmm_amd import workloads
    import os

    probs_nm = workloads.batch_problems(3, 512)
    table = os.path.join(os.path.dirname(uvt.__file__), "tune", "f64_tiles.json")
    bufs, results, errs, infos = [], {}, {}, {}
    for k, (n, m, prec) in enumerate(probs_nm):
        lda, ldb = n + (n & 1), m + (m & 1)
        dA, dB = gpu.alloc(n * lda * 8), gpu.alloc(n * ldb * 8)
        gpu.fill_seeded(n, m, 3 + k, 1, dA, lda, dB, ldb)
        bufs.append((dA, dB, lda, ldb))
    for route in ("auto", "int8"):
        outs, probs = [], []
        for k, (n, m, prec) in enumerate(probs_nm):
            dA, dB, lda, ldb = bufs[k]
            dC = gpu.alloc(n * m * 8)
            outs.append(dC)
            r = uvt.ROUTE_INT8 if (route == "int8" and prec == uvt.PREC_SPLIT) else uvt.ROUTE_AUTO
            probs.append(uvt.Problem(dA.ptr, lda, dB.ptr, ldb, dC.ptr, m, n, m, prec, r))
        batch = uvt.Batch(gpu, probs, slices=12, table=table)
        infos[route] = batch.info()
        batch.run()
        gpu.sync()
        errs[route] = batch.errors()
        batch.close()
        results[route] = [outs[k].download((n, m)) for k, (n, m, _) in enumerate(probs_nm)]
        for b in outs:
            b.free()
    say("auto", infos["auto"], "int8", infos["int8"])
    assert infos["auto"]["int8_flops"] == 0.0
    assert infos["int8"]["int8_flops"] > 0.4 * infos["int8"]["flops"]
    for k, (n, m, prec) in enumerate(probs_nm):
        if prec == uvt.PREC_F64:
            np.testing.assert_array_equal(results["int8"][k], results["auto"][k], err_msg=str(k))
        else:
            assert rel_fro(results["int8"][k], results["auto"][k]) <= 1e-12, (k, n, m)
            assert (errs["int8"][k][0] >= 0) == (n > 256)
    order = sorted(range(512), key=lambda k: -probs_nm[k][1] * probs_nm[k][0] ** 2)
    rng = np.random.default_rng(5)
    sample = sorted(set(order[:12]) | set(int(x) for x in rng.choice(512, 12, replace=False)))
    for k in sample:
        n, m, prec = probs_nm[k]
        ref = oracle.trmm(oracle.seeded_A(n, 3 + k, 1), oracle.seeded_B(n, m, 3 + k, 1))
        for route in ("auto", "int8"):
            err = rel_fro(results[route][k], ref)
            tol = 1e-13 if (prec == uvt.PREC_F64 or route == "auto") else 1e-12
            assert err <= tol, (k, n, m, prec, route, err)
            bound = errs[route][k][0]
            if bound >= 0:
                assert err <= bound, (k, n, m, err, bound)
    for dA, dB, _, _ in bufs:
        dA.free()
        dB.free()
