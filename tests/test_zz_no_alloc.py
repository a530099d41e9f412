"""uvt_dtrmm_noalloc, the allocation-free form of uvt_dtrmm (its workspace
is the library's fixed per-device scratch array instead of the handle's
growable workspace): through the C ABI it must give the same bits as
uvt_dtrmm on the same inputs, call after call (the second call smaller than
the first, so stale scratch contents from the first must not leak into the
second), for every packing path (left/right, lower/upper, transposed, unit),
and refuse a call whose workspace exceeds the scratch with UVT_ERR_SIZE."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LEFT, RIGHT = 0, 1
LOWER, UPPER = 0, 1
NOTRANS, TRANS = 0, 1
NONUNIT, UNIT = 0, 1
c = ctypes


def say(*a):
    print("[gpu-noalloc]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L(uvt, gpu):
    lib = uvt.lib()
    sig = [c.c_void_p, c.c_int, c.c_int, c.c_int, c.c_int, c.c_int64, c.c_int64, c.c_double,
           c.c_void_p, c.c_int64, c.c_void_p, c.c_int64, c.c_void_p]
    for name in ("uvt_dtrmm", "uvt_dtrmm_noalloc"):
        getattr(lib, name).argtypes = sig
        getattr(lib, name).restype = c.c_int
    return lib


def run(L, gpu, fn, side, uplo, ta, diag, A, B, alpha):
    m, n = B.shape
    dA, dB = gpu.array(A), gpu.array(B)
    try:
        st = getattr(L, fn)(gpu.ptr, side, uplo, ta, diag, m, n, alpha, dA.ptr, A.shape[1], dB.ptr,
                            n, gpu.stream)
        gpu.sync()
        return st, dB.download((m, n))
    finally:
        dA.free()
        dB.free()


FORMS = [(LEFT, LOWER, NOTRANS, NONUNIT), (LEFT, UPPER, TRANS, UNIT), (RIGHT, LOWER, TRANS, NONUNIT),
         (RIGHT, UPPER, NOTRANS, UNIT), (LEFT, LOWER, TRANS, UNIT), (RIGHT, UPPER, TRANS, NONUNIT)]


@pytest.mark.parametrize("form", FORMS)
def test_noalloc_equals_dtrmm_twice(L, gpu, form):
    side, uplo, ta, diag = form
    rng = np.random.default_rng(11 + sum(form))
    for m, n in ((1100, 700), (301, 257)):  # second call smaller than the first
        nt = m if side == LEFT else n
        A = rng.uniform(-1, 1, (nt, nt))
        B = rng.uniform(-1, 1, (m, n))
        say("form", form, "m n", m, n)
        st0, ref = run(L, gpu, "uvt_dtrmm", side, uplo, ta, diag, A, B, 0.75)
        st1, got = run(L, gpu, "uvt_dtrmm_noalloc", side, uplo, ta, diag, A, B, 0.75)
        assert st0 == 0 and st1 == 0, (st0, st1)
        np.testing.assert_array_equal(got, ref)
        # and against a float64 NumPy product of the same operands
        T = np.tril(A) if uplo == LOWER else np.triu(A)
        if diag == UNIT:
            np.fill_diagonal(T, 1.0)
        op = T.T if ta == TRANS else T
        want = 0.75 * (op @ B if side == LEFT else B @ op)
        assert np.linalg.norm(got - want) / np.linalg.norm(want) <= 1e-13


def test_noalloc_refuses_more_than_its_scratch(L, gpu, uvt):
    # left, lower, no transpose, non-unit: R alone is 8 * m * n bytes; 8192 x
    # 8192 is 512 MiB, over the 128 MiB scratch (checked before any launch,
    # so the operands need not exist)
    m = n = 8192
    st = L.uvt_dtrmm_noalloc(gpu.ptr, LEFT, LOWER, NOTRANS, NONUNIT, m, n, 1.0, c.c_void_p(256),
                             m, c.c_void_p(256), n, gpu.stream)
    assert st == uvt.ERR_SIZE
    # alpha == 0 needs no scratch: B is zeroed whatever its size
    dB = gpu.array(np.ones((64, 64)))
    st = L.uvt_dtrmm_noalloc(gpu.ptr, LEFT, LOWER, NOTRANS, NONUNIT, 64, 64, 0.0, None, 64, dB.ptr,
                             64, gpu.stream)
    gpu.sync()
    assert st == 0
    assert (dB.download((64, 64)) == 0).all()
    dB.free()
