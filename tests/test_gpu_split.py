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
