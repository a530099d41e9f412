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


def exact_sample(n, count=192):
    """Rows for exact (double-double) checks: all of them for small n, else
    the first and last of every 256-row tile plus seeded others."""
    if n <= count:
        return np.arange(n)
    rows = set(range(0, n, 256)) | set(range(255, n, 256)) | {n - 1}
    rng = np.random.default_rng(n)
    rows |= set(int(x) for x in rng.integers(0, n, count))
    return np.array(sorted(rows), dtype=np.int64)


def split_with_error(gpu, A, B, slices, lda=None):
    """(C, bound, estimate) of one split call on host operands."""
    n, m = B.shape
    lda = lda or A.shape[1]
    ws = gpu.split_workspace(n, m, slices)
    dA, dB, dC, dW = gpu.array(A), gpu.array(B), gpu.alloc(n * m * 8), gpu.alloc(ws)
    try:
        gpu.trmm_split(n, m, slices, dA, lda, dB, m, dC, m, dW)
        gpu.sync()
        bound, est = gpu.split_error(n, m, slices, dW)
        return dC.download((n, m)), bound, est
    finally:
        for b in (dA, dB, dC, dW):
            b.free()


@pytest.mark.parametrize("n,m,spread,slices", [(1024, 512, 1, 12), (2048, 256, 1, 10),
                                               (3000, 300, 0, 12), (4096, 512, 1, 13),
                                               (1500, 64, 1, 8), (600, 100, 0, 13),
                                               (200, 50, 1, 12)])
def test_error_statement_against_the_oracle(gpu, oracle, n, m, spread, slices):
    """Each call's stated error: the rigorous bound is at least the error
    measured against the oracle (rel. Frobenius), and the estimate (the
    expected value under independent uniform roundings) is within a factor 4
    of it whenever the integer part's rounding dominates."""
    A = oracle.seeded_A(n, 40 + n, spread)
    B = oracle.seeded_B(n, m, 40 + n, spread)
    C, bound, est = split_with_error(gpu, A, B, slices)
    # exact rows (double-double sums): at 13 moduli the split error is below
    # the rounding error of a plain FP64 product
    rows = exact_sample(n)
    err = rel_fro(C[rows], oracle.trmm_rows_dd(A, B, rows))
    say("n m spread slices", n, m, spread, slices, "err", err, "bound", bound, "estimate", est,
        "bound/err", bound / err if err else None)
    assert np.isfinite(bound) and np.isfinite(est)
    assert err <= bound
    assert 0.25 <= err / est <= 4.0


@pytest.mark.parametrize("n,m", [(1, 1), (5, 3), (257, 1), (259, 2), (300, 3), (520, 17)])
def test_tiny_and_ragged_shapes(gpu, oracle, n, m):
    """One row, one column, a box of one row (n = 257), B narrower than one
    16-column fragment: against the oracle, with the stated bound covering
    the error."""
    A = oracle.seeded_A(n, 40 + n, 1)
    B = oracle.seeded_B(n, m, 40 + m, 1)
    say("tiny/ragged", (n, m))
    C, bound, est = split_with_error(gpu, A, B, 13)
    err = rel_fro(C, oracle.trmm(A, B))
    say("err", err, "bound", bound, "estimate", est)
    assert err <= TOL
    assert err <= bound


def test_extreme_row_scales(gpu, oracle):
    """Rows of A scaled by 2^s, s from -950 to 950 (their sums of squares
    leave the FP64 range): each row keeps the split path's relative accuracy
    (against the oracle, whose FP64 sums stay in range)."""
    n, m = 1300, 200
    A = oracle.seeded_A(n, 71, 1)
    B = oracle.seeded_B(n, m, 71, 1)
    s = np.rint(np.linspace(-950, 950, n)).astype(int)
    As = np.ldexp(A, s[:, None])
    C, bound, est = split_with_error(gpu, As, B, 13)
    ref = oracle.trmm_rows_dd(As, B, np.arange(n))  # exact rows
    assert np.isfinite(C).all()
    # norms of rows brought back to unit scale (exact powers of two)
    Cu, Ru = np.ldexp(C, -s[:, None]), np.ldexp(ref, -s[:, None])
    row_err = np.linalg.norm(Cu - Ru, axis=1) / np.linalg.norm(Ru, axis=1)
    say("max row error", row_err.max(), "bound", bound)
    assert row_err.max() <= 1e-12
    # the Frobenius error is that of the largest rows
    g = int(s.max())
    assert rel_fro(np.ldexp(C, -g), np.ldexp(ref, -g)) <= bound


def test_extreme_column_scales(gpu, oracle):
    """Columns of B scaled by 2^t, t from -950 to 950: per-column accuracy."""
    n, m = 1100, 300
    A = oracle.seeded_A(n, 72, 1)
    B = oracle.seeded_B(n, m, 72, 1)
    t = np.rint(np.linspace(950, -950, m)).astype(int)
    Bs = np.ldexp(B, t[None, :])
    C, bound, est = split_with_error(gpu, A, Bs, 13)
    ref = oracle.trmm_rows_dd(A, Bs, np.arange(n))  # exact rows
    assert np.isfinite(C).all()
    Cu, Ru = np.ldexp(C, -t[None, :]), np.ldexp(ref, -t[None, :])
    col_err = np.linalg.norm(Cu - Ru, axis=0) / np.linalg.norm(Ru, axis=0)
    say("max column error", col_err.max(), "bound", bound)
    assert col_err.max() <= 1e-12
    g = int(t.max())
    assert rel_fro(np.ldexp(C, -g), np.ldexp(ref, -g)) <= bound


@pytest.mark.parametrize("ea,eb", [(-600, -400), (500, 500), (-1000, 300)])
def test_whole_problem_far_from_one(gpu, oracle, ea, eb):
    """A scaled by 2^ea and B by 2^eb: the result, of order 2^(ea+eb), keeps
    the relative accuracy."""
    n, m = 900, 128
    A = np.ldexp(oracle.seeded_A(n, 73, 1), ea)
    B = np.ldexp(oracle.seeded_B(n, m, 73, 1), eb)
    C, bound, est = split_with_error(gpu, A, B, 12)
    ref = oracle.trmm_rows_dd(A, B, np.arange(n))  # exact rows
    err = rel_fro(np.ldexp(C, -(ea + eb)), np.ldexp(ref, -(ea + eb)))  # norms at unit scale
    say("ea eb", ea, eb, "err", err, "bound", bound, "estimate", est)
    assert err <= 1e-12
    assert err <= bound


def test_nonfinite_entries(gpu, oracle):
    """A NaN in the box part of row i0 of A, an infinity in B[k1][j1] (k1
    below the last 256-row tile): row i0 of C is NaN, column j1 is NaN in
    every row whose box reads B[k1] (row tiles past k1's), the FP64 diagonal
    blocks that read B[k1] give a non-finite value as the FP64 kernel does,
    every other entry equals the result on the same inputs with those
    entries zeroed (bit for bit), and the stated error is NaN."""
    n, m = 1200, 90
    A = oracle.seeded_A(n, 74, 1)
    B = oracle.seeded_B(n, m, 74, 1)
    i0, k0, k1, j1 = 700, 300, 400, 17
    A0, B0 = A.copy(), B.copy()
    A0[i0, k0] = 0.0
    B0[k1, j1] = 0.0
    C0, b0, e0 = split_with_error(gpu, A0, B0, 13)
    assert np.isfinite(C0).all() and np.isfinite(b0)
    A1, B1 = A0.copy(), B0.copy()
    A1[i0, k0] = np.nan
    B1[k1, j1] = np.inf
    C1, b1, e1 = split_with_error(gpu, A1, B1, 13)
    t1 = k1 // 256
    bad = np.zeros((n, m), dtype=bool)
    bad[i0, :] = True
    bad[(t1 + 1) * 256:, j1] = True
    diag = np.zeros((n, m), dtype=bool)
    # the FP64 diagonal blocks reading B[k1][j1]: every row of k1's 128-row
    # tile and below it in tile t1 (whole k panels are multiplied, so the
    # rows above k1 of that tile meet it through A's structural zeros)
    diag[(k1 // 128) * 128:(t1 + 1) * 256, j1] = True
    assert np.isnan(C1[bad]).all()
    assert (~np.isfinite(C1[diag & ~bad])).all()
    keep = ~bad & ~diag
    np.testing.assert_array_equal(C1[keep], C0[keep])
    assert np.isnan(b1) and np.isnan(e1)


def test_config3_full_size_rows(gpu, uvt, oracle):
    """The split headline at its full size, N = 32768, M = 8192, inputs
    U[-1,1) x 2^[-8,8] (the benchmark's seed), with 12 and 13 moduli, and the
    FP64 headline (balanced, trapezoid, 16 bands) on the same inputs: 34
    rows (tile and band boundaries, the 8192-row marks, the last rows, seeded
    others) against exact double-double rows of the oracle; 12 moduli must
    stay within 1e-12, FP64 within 1e-13, and each split call's stated bound
    must cover its error."""
    n, m, seed = 32768, 8192, 3
    bands = [b for b, _ in uvt.plan_bands(n, 16, "balanced")]
    rng = np.random.default_rng(2024)
    rows = sorted(set([0, 255, 256, 8191, 8192, 16383, 16384, 32511, 32512, 32767] + bands +
                      [int(x) for x in rng.integers(0, n, 8)]))
    say("config 3 rows", len(rows))
    dA, dB, dC = gpu.alloc(n * n * 8), gpu.alloc(n * m * 8), gpu.alloc(n * m * 8)
    got, stated = {}, {}
    try:
        gpu.fill_seeded(n, m, seed, 1, dA, n, dB, m)
        plan = uvt.Plan(n, m, 16, "balanced", "trapezoid")
        gpu.trmm_f64(plan, dA, n, dB, m, dC, m)
        gpu.sync()
        plan.close()
        got["fp64"] = dC.download_rows(rows, m)
        for nm in (12, 13):
            ws = gpu.split_workspace(n, m, nm)
            say("split", nm, "workspace GiB", ws / 2**30)
            dW = gpu.alloc(ws)
            try:
                gpu.trmm_split(n, m, nm, dA, n, dB, m, dC, m, dW)
                gpu.sync()
                stated[nm] = gpu.split_error(n, m, nm, dW)
                got[nm] = dC.download_rows(rows, m)
            finally:
                dW.free()
    finally:
        for b in (dA, dB, dC):
            b.free()
    ref = oracle.seeded_rows_dd(n, m, seed, 1, rows)
    err = {k: rel_fro(v, ref) for k, v in got.items()}
    say("errors on the rows", err, "stated (bound, estimate)", stated)
    assert err["fp64"] <= 1e-13
    assert err[12] <= 1e-12
    assert err[13] <= 1e-12
    for nm in (12, 13):
        bound, est = stated[nm]
        # the rows are a sample: compare with some slack for the sample's spread
        assert err[nm] <= 2.0 * bound
        assert 0.25 <= err[nm] / est <= 4.0
