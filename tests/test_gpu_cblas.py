"""libuvt_cblas, the CBLAS drop-in for the reference's two MKL calls
(UVaTRMM/src/mm_triangular.c:347-360, :374-387): every layout / side /
triangle / transpose / diagonal of dtrmm and every transpose and alpha/beta
form of dgemm against a float64 NumPy product of the same operands, with
the unreferenced triangle (and a unit diagonal) poisoned with NaN and the
padding past each leading dimension checked untouched; then the reference's
stage loop, as the oracle restates it, calling the shim in place of MKL."""
import ctypes
import itertools
import sys

import numpy as np
import pytest

from conftest import rel_fro

pytestmark = pytest.mark.gpu

ROW, COL = 101, 102
NOTRANS, TRANS, CONJ = 111, 112, 113
UPPER, LOWER = 121, 122
NONUNIT, UNIT = 131, 132
LEFT, RIGHT = 141, 142
PAD = -12345.0
d = ctypes.c_double


def say(*a):
    print("[gpu-cblas]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cblas(uvt, gpu):
    L = ctypes.CDLL(uvt.CBLAS_PATH)
    L.cblas_dgemm.restype = None
    L.cblas_dtrmm.restype = None
    return L


def host_mat(rows, cols, ld, layout):
    """Buffer of a rows x cols matrix with leading dimension ld in the given
    layout, padding filled with PAD, and the logical view into it."""
    if layout == ROW:
        buf = np.full((max(rows, 1), ld), PAD)
        return buf, buf[:rows, :cols]
    buf = np.full((max(cols, 1), ld), PAD)
    return buf, buf[:cols, :rows].T


def pad_intact(buf, rows, cols, layout):
    if layout == ROW:
        return (buf[:rows, cols:] == PAD).all()
    return (buf[:cols, rows:] == PAD).all()


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.mark.parametrize("layout,side,uplo,trans,diag",
                         list(itertools.product((ROW, COL), (LEFT, RIGHT), (UPPER, LOWER),
                                                (NOTRANS, TRANS, CONJ), (NONUNIT, UNIT))))
def test_dtrmm_variants(cblas, layout, side, uplo, trans, diag):
    rng = np.random.default_rng(hash((layout, side, uplo, trans, diag)) % 2**32)
    for m, n in ((70, 45), (1, 3), (131, 129)):
        nt = m if side == LEFT else n
        lda, ldb = nt + 3, (n if layout == ROW else m) + 5
        abuf, A = host_mat(nt, nt, lda, layout)
        bbuf, B = host_mat(m, n, ldb, layout)
        full = rng.uniform(-1, 1, (nt, nt)) + 2.0 * np.eye(nt)
        T = np.tril(full) if uplo == LOWER else np.triu(full)
        A[...] = np.nan  # the other triangle must not be read
        mask = np.tril(np.ones((nt, nt), bool)) if uplo == LOWER else \
            np.triu(np.ones((nt, nt), bool))
        A[mask] = T[mask]
        if diag == UNIT:
            np.fill_diagonal(A, np.nan)
            np.fill_diagonal(T, 1.0)
        B0 = rng.uniform(-1, 1, (m, n))
        B[...] = B0
        opT = T if trans == NOTRANS else T.T
        alpha = -0.75
        ref = alpha * (opT @ B0 if side == LEFT else B0 @ opT)
        cblas.cblas_dtrmm(layout, side, uplo, trans, diag, m, n, d(alpha), ptr(abuf), lda,
                          ptr(bbuf), ldb)
        err = rel_fro(np.array(B), ref)
        assert err <= 1e-14, (m, n, err)
        assert pad_intact(bbuf, m, n, layout)


@pytest.mark.parametrize("layout,ta,tb", list(itertools.product((ROW, COL), (NOTRANS, TRANS, CONJ),
                                                               (NOTRANS, TRANS))))
def test_dgemm_variants(cblas, layout, ta, tb):
    rng = np.random.default_rng(ta * 7 + tb * 3 + layout)
    for (m, n, k), alpha, beta in (((65, 47, 93), 1.0, 0.0), ((65, 47, 93), -1.5, 0.5),
                                   ((300, 260, 129), 0.25, 0.0), ((5, 3, 1), 2.0, -1.0)):
        ra, ca = (k, m) if ta != NOTRANS else (m, k)
        rb, cb = (n, k) if tb != NOTRANS else (k, n)
        lda = (ca if layout == ROW else ra) + 2
        ldb = (cb if layout == ROW else rb) + 1
        ldc = (n if layout == ROW else m) + 3
        abuf, A = host_mat(ra, ca, lda, layout)
        bbuf, B = host_mat(rb, cb, ldb, layout)
        cbuf, C = host_mat(m, n, ldc, layout)
        A[...] = rng.uniform(-1, 1, (ra, ca))
        B[...] = rng.uniform(-1, 1, (rb, cb))
        C0 = rng.uniform(-1, 1, (m, n)) if beta != 0.0 else np.full((m, n), np.nan)
        C[...] = C0
        opA = np.array(A) if ta == NOTRANS else np.array(A).T
        opB = np.array(B) if tb == NOTRANS else np.array(B).T
        ref = alpha * (opA @ opB) + (beta * C0 if beta != 0.0 else 0.0)
        cblas.cblas_dgemm(layout, ta, tb, m, n, k, d(alpha), ptr(abuf), lda, ptr(bbuf), ldb,
                          d(beta), ptr(cbuf), ldc)
        err = rel_fro(np.array(C), ref)
        assert err <= 1e-14, ((m, n, k), alpha, beta, err)
        assert pad_intact(cbuf, m, n, layout)


def test_degenerate_forms(cblas):
    """k = 0 and alpha = 0 give C = beta C (A and B not read: NULL here);
    alpha = 0 in dtrmm zeroes B."""
    C = np.full((4, 6), 3.0)
    cblas.cblas_dgemm(ROW, NOTRANS, NOTRANS, 4, 6, 0, d(1.0), None, 1, None, 6, d(0.5), ptr(C), 6)
    assert (C == 1.5).all()
    cblas.cblas_dgemm(ROW, NOTRANS, NOTRANS, 4, 6, 5, d(0.0), None, 5, None, 6, d(0.0), ptr(C), 6)
    assert (C == 0.0).all()
    B = np.full((3, 4), 2.0)
    cblas.cblas_dtrmm(ROW, LEFT, LOWER, NOTRANS, NONUNIT, 3, 4, d(0.0), None, 3, ptr(B), 4)
    assert (B == 0.0).all()


@pytest.mark.parametrize("n,m,p,mapping", [(1000, 300, 4, "reg"), (777, 130, 5, "balanced"),
                                           (2048, 96, 16, "balanced")])
def test_reference_stage_loop_through_shim(uvt, oracle, n, m, p, mapping):
    """The reference's per-band calls (dgemm box, then dtrmm of the diagonal
    triangle on a saved copy of B, then C += B'), with their own pointer
    offsets and leading dimensions, made against libuvt_cblas."""
    A = oracle.seeded_A(n, 41, 1)
    B = oracle.seeded_B(n, m, 41, 1)
    ref = oracle.trmm(A, B)
    assert oracle.L.orc_blas_use(uvt.CBLAS_PATH.encode()) == 1
    try:
        rc, C = oracle.stage_loop(A, B, p, mapping, blas=1)
    finally:
        oracle.L.orc_blas_use(None)
    assert rc == 0
    err = rel_fro(C, ref)
    say(n, m, p, mapping, "rel_fro", err)
    assert err <= 1e-13
