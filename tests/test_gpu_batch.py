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
