"""Split-precision TRMM (int8 modular products + FP64 diagonal blocks) on the
MI355X, against the FP64 oracle.

Tolerance: relative Frobenius error <= 1e-12 with the default 13 moduli (the
bar stated for the split path); exact equality on integer inputs, where
every step of the method is exact.
"""
import sys

import numpy as np
import pytest

from conftest import rel_fro

pytestmark = pytest.mark.gpu
TOL = 1e-12


def say(*a):
    print("[gpu-split]", *a, flush=True)
    sys.stdout.flush()


def run_split(gpu, A, B, slices=13, lda=None):
    n, m = B.shape
    lda = lda or A.shape[1]
    say("trmm_split", dict(n=n, m=m, slices=slices, lda=lda))
    ws = gpu.split_workspace(n, m, slices)
    dA, dB, dC, dW = gpu.array(A), gpu.array(B), gpu.alloc(n * m * 8), gpu.alloc(ws)
    try:
        gpu.trmm_split(n, m, slices, dA, lda, dB, m, dC, m, dW)
        gpu.sync()
        return dC.download((n, m))
    finally:
        for b in (dA, dB, dC, dW):
            b.free()


def test_exact_on_integer_inputs(gpu, oracle):
    """Small integers: scaling, residues, int8 products and the CRT are all
    exact, so the result must equal the exact product bit for bit."""
    rng = np.random.default_rng(3)
    for n, m in ((600, 300), (1000, 257), (513, 700)):
        A = np.tril(rng.integers(-9, 10, (n, n))).astype(np.float64)
        B = rng.integers(-9, 10, (n, m)).astype(np.float64)
        C = run_split(gpu, A, B)
        np.testing.assert_array_equal(C, oracle.trmm(A, B))


@pytest.mark.parametrize("n,m,spread", [(1024, 512, 0), (1500, 333, 1), (2049, 1000, 1),
                                        (300, 64, 1), (200, 50, 0), (4096, 256, 1)])
def test_error_vs_fp64_oracle(gpu, oracle, n, m, spread):
    A = oracle.seeded_A(n, 11 + n, spread)
    B = oracle.seeded_B(n, m, 11 + n, spread)
    err = rel_fro(run_split(gpu, A, B), oracle.trmm(A, B))
    say("rel_fro", err)
    assert err <= TOL


def test_more_moduli_smaller_error(gpu, oracle):
    n, m = 1200, 200
    A = oracle.seeded_A(n, 5, 1)
    B = oracle.seeded_B(n, m, 5, 1)
    ref = oracle.trmm(A, B)
    errs = [rel_fro(run_split(gpu, A, B, s), ref) for s in (8, 10, 13)]
    say("errors by moduli", errs)
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] <= TOL


def test_odd_leading_dimension(gpu, oracle):
    n, m = 700, 90
    A = oracle.seeded_A(n, 2, 1)
    B = oracle.seeded_B(n, m, 2, 1)
    Abig = np.zeros((n, n + 1))
    Abig[:, :n] = A
    assert rel_fro(run_split(gpu, Abig, B, lda=n + 1), oracle.trmm(A, B)) <= TOL


def test_workspace_errors(gpu, uvt):
    with pytest.raises(uvt.UvtError):
        gpu.split_workspace(100, 100, 1)
    with pytest.raises(uvt.UvtError):
        gpu.split_workspace(100, 100, 17)


def test_twelve_moduli_large_k(gpu, uvt, oracle):
    """12 moduli at a deep reduction (n = 16384, exponents over 2^[-8, 8]) stay
    within the 1e-12 bar; the reference is the FP64 path (itself within
    1e-13 of the oracle), here computed on the GPU."""
    n, m = 16384, 96
    say("12 moduli", n, m)
    dA, dB = gpu.alloc(n * n * 8), gpu.alloc(n * m * 8)
    dC, dS = gpu.alloc(n * m * 8), gpu.alloc(n * m * 8)
    ws = gpu.split_workspace(n, m, 12)
    dW = gpu.alloc(ws)
    try:
        gpu.fill_seeded(n, m, 77, 1, dA, n, dB, m)
        plan = uvt.Plan(n, m, 8, "balanced", "trapezoid")
        gpu.trmm_f64(plan, dA, n, dB, m, dC, m)
        gpu.trmm_split(n, m, 12, dA, n, dB, m, dS, m, dW)
        gpu.sync()
        ref, got = dC.download((n, m)), dS.download((n, m))
        plan.close()
    finally:
        for b in (dA, dB, dC, dS, dW):
            b.free()
    err = rel_fro(got, ref)
    say("rel_fro 12 moduli", err)
    assert err <= TOL



def test_max_order_65536_split(gpu, oracle):
    """Split path at N = 65536 (reduction depth 65280, the int32 accumulation
    bound's range) with 13 moduli: sampled rows against double-double sums."""
    n, m = 65536, 64
    say("split max order", n, m)
    ws = gpu.split_workspace(n, m, 13)
    dA, dB, dC, dW = gpu.alloc(n * n * 8), gpu.alloc(n * m * 8), gpu.alloc(n * m * 8), gpu.alloc(ws)
    try:
        gpu.fill_seeded(n, m, 92, 1, dA, n, dB, m)
        gpu.trmm_split(n, m, 13, dA, n, dB, m, dC, m, dW)
        gpu.sync()
        C = dC.download((n, m))
    finally:
        for b in (dA, dB, dC, dW):
            b.free()
    rows = [0, 255, 256, 257, 30000, 65279, 65280, 65535]
    ref = oracle.seeded_rows_dd(n, m, 92, 1, rows)
    got = C[rows]
    err = rel_fro(got, ref)
    say("rel_fro split 65536", err)
    assert err <= TOL
