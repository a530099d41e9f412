"""Host <-> device copies (uvt_memcpy, uvt_memcpy_2d): contiguous copies of
1 MiB or more and strided copies go through the handle's two 16 MiB pinned
buffers a chunk at a time; every byte arrives, in several chunks and at
chunk edges, and nothing outside the rows is written.  Then the CBLAS
drop-in on operands large enough to take several chunks, with small-integer
entries (exact in any summation order)."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHUNK = 16 << 20


def say(*a):
    print("[gpu-copies]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("nbytes", [8, (1 << 20) - 8, 1 << 20, CHUNK, CHUNK + 8, 3 * CHUNK + 4096 + 8])
def test_contiguous_round_trip(gpu, nbytes):
    rng = np.random.default_rng(nbytes)
    a = rng.integers(0, 2**63, nbytes // 8, dtype=np.int64).view(np.float64)
    d = gpu.alloc(nbytes)
    try:
        d.upload(a)
        np.testing.assert_array_equal(d.download(a.shape).view(np.int64), a.view(np.int64))
    finally:
        d.free()


# (rows, width bytes, host pitch, device pitch): several chunks with packed
# and pitched device rows, a chunk edge inside the rows, one row, and rows
# wider than a chunk
@pytest.mark.parametrize("rows,width,hp,dp", [
    (3000, 8000, 8800, 8000),
    (3000, 8000, 8800, 8016),
    (2097, 8008, 8008 + 24, 8016),
    (1, 4096, 8192, 4096),
    (3, CHUNK + 8, CHUNK + 64, CHUNK + 16),
])
def test_strided_round_trip(gpu, rows, width, hp, dp):
    rng = np.random.default_rng(rows + width)
    host = rng.integers(0, 255, (rows - 1) * hp + width, dtype=np.uint8)
    d = gpu.alloc((rows - 1) * dp + width)
    back = np.full((rows - 1) * hp + width + 64, 7, dtype=np.uint8)  # pitch hp, 64 spare bytes
    try:
        d.zero()
        d.upload_2d(host, width, rows, dp, hp)
        dev = d.download(((rows - 1) * dp + width,), np.uint8)
        d.download_2d(back, width, rows, hp, dp)
    finally:
        d.free()
    for r in range(rows):
        np.testing.assert_array_equal(dev[r * dp:r * dp + width], host[r * hp:r * hp + width])
        if r + 1 < rows:  # device padding: never written
            assert (dev[r * dp + width:(r + 1) * dp] == 0).all()
            assert (back[r * hp + width:(r + 1) * hp] == 7).all()
        np.testing.assert_array_equal(back[r * hp:r * hp + width], host[r * hp:r * hp + width])
    assert (back[(rows - 1) * hp + width:] == 7).all()


def test_cblas_dtrmm_large_strided(uvt, gpu):
    """cblas_dtrmm, row-major left lower, on a 3000 x 3000 A with lda 3001
    and a 3000 x 1500 B with ldb 1503 (72 and 36 MB: several chunks each
    way), against NumPy; the padding of B is left as it was."""
    L = ctypes.CDLL(uvt.CBLAS_PATH)
    L.cblas_dtrmm.restype = None
    n, m, lda, ldb = 3000, 1500, 3001, 1503
    rng = np.random.default_rng(5)
    A = np.full((n, lda), np.nan)
    A[:, :n] = np.where(np.tril(np.ones((n, n), bool)), rng.integers(-3, 4, (n, n)), np.nan)
    B = np.full((n, ldb), -7.0)
    B[:, :m] = rng.integers(-3, 4, (n, m))
    want = np.tril(np.nan_to_num(A[:, :n])) @ B[:, :m]
    L.cblas_dtrmm(101, 141, 122, 111, 131, n, m, ctypes.c_double(1.0),
                  A.ctypes.data_as(ctypes.c_void_p), lda, B.ctypes.data_as(ctypes.c_void_p), ldb)
    np.testing.assert_array_equal(B[:, :m], want)
    assert (B[:, m:] == -7.0).all()
