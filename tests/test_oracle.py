"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

Sources of truth, none of which is the oracle's own formula retyped:
* band vectors and abort codes produced by the reference's own mapping
  (tests/golden/mapping_golden.json);
* the reference program's printed Norm values (tests/golden/reference_norms.json);
* MKL cblas_dtrmm / the reference's dgemm+dtrmm stage loop run with MKL
  (tests/golden/mkl_*.npz);
* published SplitMix64 outputs; OpenBLAS via numpy/scipy; exact rational
  arithmetic on tiny inputs; algebraic identities.
"""
import json
import os
import re
import sys
from fractions import Fraction

import numpy as np
import pytest

from conftest import GOLDEN, rel_fro

sys.path.insert(0, GOLDEN)
import make_golden as gold  # noqa: E402  (numpy restatement of the input generators)


def _golden_mapping():
    with open(os.path.join(GOLDEN, "mapping_golden.json")) as f:
        return json.load(f)


def _parse(bands):
    return [tuple(int(x) for x in b.split(":")) for b in bands.split(",")]


@pytest.mark.parametrize("case", _golden_mapping()["bands"], ids=lambda c: f"{c['n']}-{c['p']}-{c['map']}")
def test_mapping_golden(oracle, case):
    rc, bands = oracle.bands(case["n"], case["p"], case["map"])
    assert rc == 0
    assert bands == _parse(case["bands"])


def test_mapping_abort_codes(oracle):
    rc, _ = oracle.bands(7, 7, "balanced")
    assert rc == 11  # check_range -> INVALID_RANGE + 2
    rc, _ = oracle.map(1, 4, 0, "balanced")  # 1 element over 4 bands -> ideal < 1
    assert rc == 7
    rc, _ = oracle.map(100000, 4, 0, "balanced")  # 32-bit count overflow: undefined in the reference
    assert rc == -1


def test_mapping_partitions_rows(oracle):
    """Bands tile [0, n) in order (a property of both mappings)."""
    for n in (1, 2, 3, 10, 97, 1000, 2048, 4097):
        for p in (1, 2, 3, 4, 7, 8, 16):
            if p > n:
                continue
            for mp in ("reg", "balanced"):
                rc, bands = oracle.bands(n, p, mp)
                if rc:
                    assert mp == "balanced" and rc == 11
                    continue
                pos = 0
                for b, s in bands:
                    assert b == pos and s > 0
                    pos += s
                assert pos == n


def test_balanced_equalises_nonzeros(oracle):
    """Balanced bands hold near-equal lower-triangle element counts."""
    n, p = 16384, 16
    _, bands = oracle.bands(n, p, "balanced")
    counts = [sum(range(b + 1, b + s + 1)) for b, s in bands]
    ideal = n * (n + 1) / 2 / p
    assert max(abs(c - ideal) for c in counts[:-1]) / ideal < 0.01


def test_splitmix64_known_answers(oracle):
    # Published SplitMix64 outputs for state 1234567 (first five values).
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423,
            4593380528125082431, 16408922859458223821]
    assert [oracle.L.orc_splitmix64(1234567, k) for k in range(1, 6)] == want


def test_seeded_generator_matches_numpy_restatement(oracle):
    for spread in (0, 1):
        A = oracle.seeded_A(37, 11, spread)
        B = oracle.seeded_B(37, 29, 11, spread)
        np.testing.assert_array_equal(A, gold.seeded_A(37, 11, spread))
        np.testing.assert_array_equal(B, gold.seeded_B(37, 29, 11, spread))
    A = oracle.seeded_A(64, 3, 0)
    off = A[np.tril_indices(64, -1)]
    assert off.min() >= -1 and off.max() < 1 and abs(off.mean()) < 0.1
    assert np.all(np.diag(A) >= 1) and np.all(np.diag(A) < 3)
    As = oracle.seeded_A(200, 5, 1)
    e = np.frexp(As[np.tril_indices(200, -1)])[1]
    assert e.min() <= -8 and e.max() >= 8  # spread reaches 2^[-8, 8]


def test_closed_form_elements(oracle):
    A = oracle.closed_A(50)
    B = oracle.closed_B(50, 70)
    assert A[0, 0] == 1.0 and A[3, 4] == 0.0
    assert A[10, 3] == 1.0 / (10 * 50 + 3 + 1)
    assert B[49, 69] == 1.0 / ((49 * 50 + 69) % 2341 + 1)
    np.testing.assert_array_equal(A, gold.closed_A(50))
    np.testing.assert_array_equal(B, gold.closed_B(50, 70))


@pytest.mark.parametrize("n", [257, 1000, 2048])
def test_reference_norms(oracle, n):
    """The reference program's printed Norm on its closed-form inputs."""
    with open(os.path.join(GOLDEN, "reference_norms.json")) as f:
        norms = {d["n"]: d["norm"] for d in json.load(f)["norms"]}
    C = oracle.trmm(oracle.closed_A(n), oracle.closed_B(n, n))
    v, _ = oracle.inf_norm(C)
    assert "%.3f" % v == norms[n]


def test_trmm_matches_mkl_fixtures(oracle):
    d = np.load(os.path.join(GOLDEN, "mkl_trmm.npz"))
    keys = [k[:-2] for k in d.files if k.endswith("_C")]
    assert len(keys) >= 4
    for key in keys:
        seed, n, m, sp = (int(v) for v in re.fullmatch(r"s(\d+)_n(\d+)_m(\d+)_sp(\d+)", key).groups())
        A = oracle.seeded_A(n, seed, sp)
        B = oracle.seeded_B(n, m, seed, sp)
        assert A[0, 0] == d[key + "_A00"] and np.sum(B) == d[key + "_Bsum"]
        C = oracle.trmm(A, B)
        assert rel_fro(C, d[key + "_C"]) < 1e-14, key


def test_stage_loop_matches_mkl_fixtures(oracle):
    d = np.load(os.path.join(GOLDEN, "mkl_stage_loop.npz"))
    for key in d.files:
        if not key.endswith("_C"):
            continue
        n, p, mp, _ = key.split("_")
        n, p = int(n[1:]), int(p[1:])
        A, B = oracle.closed_A(n), oracle.closed_B(n, n)
        rc, C = oracle.stage_loop(A, B, p, mp, blas=0)
        assert rc == 0
        assert rel_fro(C, d[key]) < 1e-14, key
        assert rel_fro(oracle.trmm(A, B), d[key]) < 1e-14, key


def test_trmm_against_openblas(oracle):
    from scipy.linalg import blas
    A = oracle.seeded_A(150, 9)
    B = oracle.seeded_B(150, 77, 9)
    # scipy's dtrmm is column-major: C^T = B^T tril(A)^T, i.e. right side, upper, transposed.
    ref = blas.dtrmm(1.0, A, B, lower=1)
    assert rel_fro(oracle.trmm(A, B), ref) < 1e-14
    assert rel_fro(oracle.trmm(A, B), np.tril(A) @ B) < 1e-14


def test_trmm_exact_on_tiny_inputs(oracle):
    """Exact rational arithmetic bounds the oracle's rounding on small cases."""
    rng = np.random.default_rng(0)
    for n, m in ((1, 1), (3, 2), (7, 5), (12, 9)):
        A = np.tril(rng.integers(-8, 9, (n, n)) / 8.0)
        B = rng.integers(-8, 9, (n, m)) / 4.0
        C = oracle.trmm(A, B)  # dyadic inputs, small sums: exact in double
        for i in range(n):
            for j in range(m):
                exact = sum(Fraction(A[i, k]) * Fraction(B[k, j]) for k in range(i + 1))
                assert Fraction(C[i, j]) == exact


def test_trmm_identities(oracle):
    n, m = 64, 33
    B = oracle.seeded_B(n, m, 4)
    np.testing.assert_array_equal(oracle.trmm(np.eye(n), B), B)
    L1 = np.tril(np.ones((n, n)))
    np.testing.assert_allclose(oracle.trmm(L1, B), np.cumsum(B, axis=0), rtol=0, atol=1e-13)
    # the upper triangle of A must be ignored
    A = oracle.seeded_A(n, 4)
    Afull = A + np.triu(np.full((n, n), 7.0), 1)
    np.testing.assert_array_equal(oracle.trmm(Afull, B), oracle.trmm(A, B))
    # linearity in B
    B2 = oracle.seeded_B(n, m, 5)
    np.testing.assert_allclose(oracle.trmm(A, B + B2), oracle.trmm(A, B) + oracle.trmm(A, B2),
                               rtol=0, atol=1e-13)


def test_rows_dd_is_accurate(oracle):
    n, m = 300, 40
    A = oracle.seeded_A(n, 21, 1)
    B = oracle.seeded_B(n, m, 21, 1)
    rows = np.array([0, 1, 150, 299])
    dd = oracle.trmm_rows_dd(A, B, rows)
    for r, i in enumerate(rows):
        for j in (0, 17, 39):
            exact = sum(Fraction(A[i, k]) * Fraction(B[k, j]) for k in range(i + 1))
            assert abs(Fraction(dd[r, j]) - exact) <= abs(exact) * Fraction(1, 2**52)
    assert rel_fro(dd, oracle.trmm(A, B)[rows]) < 1e-14


def test_inf_norm_matches_numpy(oracle):
    C = oracle.seeded_B(70, 45, 8, 1)
    v, rs = oracle.inf_norm(C)
    assert v == np.linalg.norm(C, np.inf)
    np.testing.assert_allclose(rs, np.abs(C).sum(axis=1), rtol=1e-15)


def test_stage_loop_blas_path(oracle):
    """The MKL-backed replay (CPU baseline) equals the C-loop replay."""
    import ctypes
    buf = ctypes.create_string_buffer(256)
    if not oracle.L.orc_blas_available(buf, 256):
        pytest.skip("MKL not installed on this host")
    A, B = oracle.closed_A(97), oracle.closed_B(97, 97)
    rc0, C0 = oracle.stage_loop(A, B, 3, "balanced", blas=0)
    rc1, C1 = oracle.stage_loop(A, B, 3, "balanced", blas=1)
    assert rc0 == rc1 == 0
    assert rel_fro(C1, C0) < 1e-14


def test_seeded_rows_dd_equals_rows_dd_over_filled_inputs(oracle):
    """The on-the-fly row oracle (used at orders whose A does not fit in host
    memory) equals orc_trmm_rows_dd over orc_fill_seeded_A/B bit for bit:
    same elements, same double-double sums in the same order."""
    n, m = 700, 37
    rows = [0, 1, 255, 256, 699]
    for spread in (0, 1):
        A = oracle.seeded_A(n, 9, spread)
        B = oracle.seeded_B(n, m, 9, spread)
        want = oracle.trmm_rows_dd(A, B, rows)
        got = oracle.seeded_rows_dd(n, m, 9, spread, rows)
        np.testing.assert_array_equal(got, want)
