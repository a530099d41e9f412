"""Writes the golden fixtures under tests/golden/.

Run once in a container that has Intel MKL (libmkl_rt) installed:

    python tests/golden/make_golden.py

Fixtures written:

* ``mapping_golden.json`` -- band vectors (begin:size) and abort exit codes
  produced by the reference's own ``map_regular`` / ``map_balanced``
  (UVaTRMM/src/mm_mapping.c:27-67, 319-366) and by the reference program's
  argument checks, as recorded when the reference was built and run from its
  sources.  Stored as data; nothing here re-runs the reference.
* ``reference_norms.json`` -- the ``Norm: %.3lf`` line printed by the
  reference program (UVaTRMM/src/mm_triangular_common.c:2987) for
  ``uvatrmm.x N <map> <shape>`` on its closed-form inputs.
* ``mkl_trmm.npz`` -- C = tril(A) * B computed by MKL ``cblas_dtrmm``
  (RowMajor, Left, Lower, NoTrans, NonUnit, alpha=1), the BLAS the reference
  links (UVaTRMM/Makefile:19), on seeded inputs generated here in numpy.
* ``mkl_stage_loop.npz`` -- the reference's per-rank stage loop
  (UVaTRMM/src/mm_triangular.c:319-410: cblas_dgemm box with K=b+1, then
  cblas_dtrmm of the (h-1) diagonal triangle and C += B) executed with MKL on
  the reference's closed-form inputs, for a few (N, P, map) cases.

The seeded generator is restated here in numpy (independently of the C
oracle), so the fixtures also pin the oracle's generator.
"""
import ctypes
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GAMMA = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def seeded(seed, stream, idx, spread, diag):
    """SplitMix64 counter generator (project-defined; see oracle/oracle.c)."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = _mix64(np.uint64(seed) ^ (np.uint64(stream) * GAMMA))
        z = _mix64(key + (idx + np.uint64(1)) * GAMMA)
    u = (z >> np.uint64(11)).astype(np.float64) * 2.0**-53
    v = 2.0 * u - 1.0
    v = np.where(diag, v + 2.0, v)
    if spread:
        e = (_mix64(z) % np.uint64(17)).astype(np.int64) - 8
        v = v * np.ldexp(1.0, e)
    return v


def seeded_A(n, seed, spread=0):
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    v = seeded(seed, 1, (i * n + j).astype(np.uint64), spread, i == j)
    return np.where(j > i, 0.0, v)


def seeded_B(n, m, seed, spread=0):
    i, j = np.meshgrid(np.arange(n), np.arange(m), indexing="ij")
    return seeded(seed, 2, (i * m + j).astype(np.uint64), spread, np.zeros_like(i, dtype=bool))


def closed_A(n):
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    v = 1.0 / (((i * n + j) % 2341) + 1).astype(np.float64)
    return np.where(j > i, 0.0, v)


def closed_B(n, m):
    i, j = np.meshgrid(np.arange(n), np.arange(m), indexing="ij")
    return 1.0 / (((i * n + j) % 2341) + 1).astype(np.float64)


# --- golden data recorded from the reference (see module docstring) -------
MAPPING = {
    "bands": [
        {"n": 2048, "p": 4, "map": "reg", "bands": "0:512,512:512,1024:512,1536:512"},
        {"n": 2048, "p": 4, "map": "balanced", "bands": "0:1023,1023:425,1448:325,1773:275"},
        {"n": 16384, "p": 16, "map": "balanced",
         "bands": "0:4095,4095:1697,5792:1302,7094:1097,8191:967,9158:874,10032:804,10836:749,"
                  "11585:702,12287:665,12952:632,13584:604,14188:580,14768:557,15325:538,15863:521"},
        {"n": 32768, "p": 16, "map": "balanced",
         "bands": "0:8191,8191:3393,11584:2604,14188:2195,16383:1934,18317:1749,20066:1607,"
                  "21673:1497,23170:1405,24575:1330,25905:1264,27169:1208,28377:1159,29536:1115,"
                  "30651:1076,31727:1041"},
        {"n": 32768, "p": 4, "map": "balanced", "bands": "0:16383,16383:6787,23170:5207,28377:4391"},
        {"n": 32768, "p": 4, "map": "reg", "bands": "0:8192,8192:8192,16384:8192,24576:8192"},
        {"n": 1000, "p": 3, "map": "reg", "bands": "0:333,333:333,666:334"},
        {"n": 10, "p": 4, "map": "balanced", "bands": "0:4,4:3,7:1,8:2"},
    ],
    "aborts": [
        {"n": 7, "p": 7, "map": "balanced", "exit": 11, "why": "check_range: a band of size 0"},
        {"argv": ["2048", "reg"], "exit": 3, "why": "argc != 4 -> show_usage"},
        {"argv": ["2048", "foo", "boxes"], "exit": 3, "why": "unknown mapping -> show_usage"},
        {"argv": ["2048", "reg", "foo"], "exit": 3, "why": "unknown shape -> show_usage"},
        {"argv": ["3", "reg", "boxes"], "p": 4, "exit": 3, "why": "N < P -> show_usage"},
    ],
}

NORMS = [  # uvatrmm.x N map shape, closed-form inputs; identical for all maps/shapes/P
    {"n": 257, "norm": "9.589"},
    {"n": 1000, "norm": "29.286"},
    {"n": 2048, "norm": "61.447"},
    {"n": 4096, "norm": "243.282"},
    {"n": 8192, "norm": "950.351"},
    {"n": 16384, "norm": "3384.807"},
    {"n": 32768, "norm": "13597.208"},
]

# --- MKL ------------------------------------------------------------------
ROW, NOT, LOW, NONUNIT, LEFT = 101, 111, 122, 131, 141


def load_mkl():
    for name in ("libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so.1"):
        try:
            lib = ctypes.CDLL(name)
            break
        except OSError:
            lib = None
    if lib is None:
        raise SystemExit("MKL (libmkl_rt) not found")
    lib.MKL_Set_Threading_Layer(1)  # sequential
    buf = ctypes.create_string_buffer(256)
    lib.MKL_Get_Version_String(buf, 256)
    d = ctypes.c_double
    p = ctypes.POINTER(ctypes.c_double)
    i = ctypes.c_int
    lib.cblas_dgemm.argtypes = [i, i, i, i, i, i, d, p, i, p, i, d, p, i]
    lib.cblas_dtrmm.argtypes = [i, i, i, i, i, i, i, d, p, i, p, i]
    return lib, buf.value.decode()


def ptr(a, off=0):
    return ctypes.cast(a.ctypes.data + 8 * off, ctypes.POINTER(ctypes.c_double))


def mkl_trmm(lib, A, B):
    n, m = B.shape
    Bc = np.ascontiguousarray(B.copy())
    lib.cblas_dtrmm(ROW, LEFT, LOW, NOT, NONUNIT, n, m, 1.0, ptr(A), n, ptr(Bc), m)
    return Bc


def map_regular(n, p, i):
    eq = max(n // p, 1)
    b = i * eq
    return b, eq + (n % p if i == p - 1 else 0)


def map_balanced(n, p, i):
    ideal = math.ceil((n * (n + 1) // 2) / p)

    def rb(k):
        num = -1 + math.sqrt(1 + 8 * (k * float(ideal)))
        return int(math.floor(num) / 2) if num > 0 else 0

    b, s = rb(i), rb(i + 1) - rb(i)
    if i == 0:
        b, s = 0, (rb(1) if p > 1 else n)
    if n > 1 and i == p - 1:
        s = n - b
    return b, s


def mkl_stage_loop(lib, A, B, p, mapping):
    """The reference's per-rank product loop (mm_triangular.c:319-410) with MKL."""
    n, w = B.shape
    C = np.zeros((n, w))
    Bw = np.ascontiguousarray(B.copy())
    for s in range(p):
        b, h = (map_balanced if mapping == "balanced" else map_regular)(n, p, s)
        saved = Bw[b:b + h].copy()
        lib.cblas_dgemm(ROW, NOT, NOT, h, w, b + 1, 1.0, ptr(A, b * n), n, ptr(Bw), w, 0.0,
                        ptr(C, b * w), w)
        tri = min(h - 1, n)
        left = b + h - tri
        if tri > 0:
            lib.cblas_dtrmm(ROW, LEFT, LOW, NOT, NONUNIT, tri, w, 1.0,
                            ptr(A, (b + h - tri) * n + left), n, ptr(Bw, left * w), w)
            C[left:left + tri] += Bw[left:left + tri]
        Bw[b:b + h] = saved
    return C


def main():
    with open(os.path.join(HERE, "mapping_golden.json"), "w") as f:
        json.dump(MAPPING, f, indent=1)
    with open(os.path.join(HERE, "reference_norms.json"), "w") as f:
        json.dump({"source": "reference program stdout 'Norm: %.3lf' on closed-form inputs",
                   "norms": NORMS}, f, indent=1)
    lib, ver = load_mkl()
    trmm_cases = [(1, 64, 48, 0), (2, 127, 33, 0), (3, 256, 128, 0), (4, 200, 64, 1), (5, 1, 7, 0)]
    out = {"mkl_version": np.array(ver)}
    for seed, n, m, spread in trmm_cases:
        A = seeded_A(n, seed, spread)
        B = seeded_B(n, m, seed, spread)
        C = mkl_trmm(lib, A, B)
        key = f"s{seed}_n{n}_m{m}_sp{spread}"
        out[key + "_C"] = C
        out[key + "_A00"] = A[0, 0]
        out[key + "_Bsum"] = np.float64(np.sum(B))
    np.savez_compressed(os.path.join(HERE, "mkl_trmm.npz"), **out)
    stage_cases = [(192, 5, "balanced"), (128, 4, "reg"), (97, 3, "balanced"), (48, 1, "reg")]
    out = {"mkl_version": np.array(ver)}
    for n, p, mp in stage_cases:
        A = closed_A(n)
        B = closed_B(n, n)
        C = mkl_stage_loop(lib, A, B, p, mp)
        out[f"n{n}_p{p}_{mp}_C"] = C
    np.savez_compressed(os.path.join(HERE, "mkl_stage_loop.npz"), **out)
    print("wrote fixtures with", ver)


if __name__ == "__main__":
    main()
