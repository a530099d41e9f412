"""Shared fixtures: the C oracle (test infrastructure) and the libuvt handle."""
import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libuvt's C ABI)")


class Oracle:
    """ctypes view of oracle/liboracle.so (the CPU checker)."""

    def __init__(self):
        if not os.path.exists(ORACLE_SO):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = ctypes.CDLL(ORACLE_SO)
        c = ctypes
        u64, i64, u32, i32 = c.c_uint64, c.c_int64, c.c_uint32, c.c_int
        P = c.POINTER(c.c_double)
        L.orc_map_regular.argtypes = [u64, u32, u32, c.POINTER(u64), c.POINTER(u64)]
        L.orc_map_balanced.argtypes = [u64, u32, u32, c.POINTER(u64), c.POINTER(u64)]
        L.orc_fill_closed_A.argtypes = [i64, P, i64]
        L.orc_fill_closed_B.argtypes = [i64, i64, P, i64]
        L.orc_seeded_value.argtypes = [u64, u32, u64, i32, i32]
        L.orc_seeded_value.restype = c.c_double
        L.orc_splitmix64.argtypes = [u64, u64]
        L.orc_splitmix64.restype = u64
        L.orc_fill_seeded_A.argtypes = [i64, u64, i32, P, i64]
        L.orc_fill_seeded_B.argtypes = [i64, i64, u64, i32, P, i64]
        L.orc_trmm.argtypes = [i64, i64, P, i64, P, i64, P, i64]
        L.orc_trmm_rows_dd.argtypes = [i64, i64, P, i64, P, i64, c.POINTER(i64), i64, P, i64]
        L.orc_stage_loop.argtypes = [i64, i64, u32, i32, P, i64, P, i64, P, i64, i32]
        L.orc_blas_available.argtypes = [c.c_char_p, i32]
        L.orc_blas_use.argtypes = [c.c_char_p]
        L.orc_seeded_rows_dd.argtypes = [i64, i64, u64, i32, c.POINTER(i64), i64, P, i64]
        L.orc_ranks_stage_loop.argtypes = [i64, i64, u32, i32, i32, P, i64, P, i64, P, i64, i32]
        L.orc_ranks_stage_loop.restype = c.c_double
        L.orc_inf_norm.argtypes = [i64, i64, P, i64, P]
        L.orc_inf_norm.restype = c.c_double
        self.L = L

    @staticmethod
    def _p(a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    def map(self, n, p, i, mapping):
        b, s = ctypes.c_uint64(), ctypes.c_uint64()
        f = self.L.orc_map_balanced if mapping == "balanced" else self.L.orc_map_regular
        rc = f(n, p, i, ctypes.byref(b), ctypes.byref(s))
        return rc, (b.value, s.value)

    def bands(self, n, p, mapping):
        out = []
        for i in range(p):
            rc, r = self.map(n, p, i, mapping)
            if rc:
                return rc, None
            out.append(r)
        return 0, out

    def closed_A(self, n):
        A = np.empty((n, n))
        self.L.orc_fill_closed_A(n, self._p(A), n)
        return A

    def closed_B(self, n, m):
        B = np.empty((n, m))
        self.L.orc_fill_closed_B(n, m, self._p(B), m)
        return B

    def seeded_A(self, n, seed, spread=0):
        A = np.empty((n, n))
        self.L.orc_fill_seeded_A(n, seed, spread, self._p(A), n)
        return A

    def seeded_B(self, n, m, seed, spread=0):
        B = np.empty((n, m))
        self.L.orc_fill_seeded_B(n, m, seed, spread, self._p(B), m)
        return B

    def trmm(self, A, B):
        n, m = B.shape
        A = np.ascontiguousarray(A)
        B = np.ascontiguousarray(B)
        C = np.empty((n, m))
        self.L.orc_trmm(n, m, self._p(A), A.shape[1], self._p(B), m, self._p(C), m)
        return C

    def trmm_rows_dd(self, A, B, rows):
        n, m = B.shape
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        out = np.empty((len(rows), m))
        self.L.orc_trmm_rows_dd(n, m, self._p(A), A.shape[1], self._p(B), m,
                                rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(rows),
                                self._p(out), m)
        return out

    def seeded_rows_dd(self, n, m, seed, spread, rows):
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        out = np.empty((len(rows), m))
        self.L.orc_seeded_rows_dd(n, m, seed, spread,
                                  rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(rows),
                                  self._p(out), m)
        return out

    def stage_loop(self, A, B, p, mapping, blas=0):
        n, m = B.shape
        C = np.zeros((n, m))
        rc = self.L.orc_stage_loop(n, m, p, 1 if mapping == "balanced" else 0, self._p(A), n,
                                   self._p(B), m, self._p(C), m, blas)
        return rc, C

    def inf_norm(self, C):
        n, m = C.shape
        rs = np.empty(n)
        v = self.L.orc_inf_norm(n, m, self._p(np.ascontiguousarray(C)), m, self._p(rs))
        return v, rs


@pytest.fixture(scope="session")
def oracle():
    return Oracle()


@pytest.fixture(scope="session")
def uvt():
    import uvatrmm_amd
    uvatrmm_amd.lib()
    return uvatrmm_amd


@pytest.fixture(scope="session")
def gpu(uvt):
    if not uvt.gpu_available():
        pytest.fail("no GPU visible to libuvt (gpu tests need an MI355X)")
    h = uvt.Handle(0)
    yield h
    h.close()


def rel_fro(x, ref):
    d = np.linalg.norm(x - ref)
    r = np.linalg.norm(ref)
    return d / r if r else d
