"""FP64 TRMM on the MI355X through libuvt's C ABI, checked against the oracle.

Tolerance: relative Frobenius error <= 1e-13 against the CPU oracle (the
bar stated for the plain FP64 path); exact equality where the arithmetic is
exact (identity / dyadic inputs, generators, repeat runs).
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN, rel_fro

pytestmark = pytest.mark.gpu
TOL = 1e-13


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def run_trmm(gpu, uvt, A, B, p=1, mapping="reg", shape="trapezoid", lda=None):
    n, m = B.shape
    lda = lda or A.shape[1]
    say("trmm_f64", dict(n=n, m=m, p=p, map=mapping, shape=shape, lda=lda))
    plan = uvt.Plan(n, m, p, mapping, shape)
    dA, dB, dC = gpu.array(A), gpu.array(B), gpu.alloc(n * m * 8)
    try:
        gpu.trmm_f64(plan, dA, lda, dB, m, dC, m)
        gpu.sync()
        return dC.download((n, m))
    finally:
        for b in (dA, dB, dC):
            b.free()
        plan.close()


def test_small_exact_cases(gpu, uvt):
    """Identity and all-ones lower A give exact results: checks the MFMA
    operand/result layouts and every tile edge."""
    for n, m in ((1, 1), (5, 3), (16, 16), (17, 33), (130, 129), (256, 300)):
        B = (np.arange(n * m, dtype=np.float64).reshape(n, m) % 7) - 3.0
        for shape, p in (("trapezoid", 1), ("boxes", min(3, n))):
            C = run_trmm(gpu, uvt, np.eye(n), B, p=p, shape=shape)
            np.testing.assert_array_equal(C, B)
            C = run_trmm(gpu, uvt, np.tril(np.ones((n, n))), B, p=p, shape=shape)
            np.testing.assert_array_equal(C, np.cumsum(B, axis=0))


INT_CASES = [  # (n, m, p, mapping, shape, tile the plan picks)
    (700, 40, 3, "reg", "boxes", (128, 32)),
    (700, 40, 1, "reg", "trapezoid", (128, 32)),
    (300, 200, 5, "balanced", "boxes", (64, 64)),
    (1001, 333, 1, "balanced", "trapezoid", (64, 64)),
    (4096, 2048, 7, "balanced", "boxes", (128, 128)),
    (4099, 2050, 1, "reg", "trapezoid", (128, 128)),
]


@pytest.mark.parametrize("n,m,p,mp,shape,tile", INT_CASES)
def test_integer_inputs_exact(gpu, uvt, n, m, p, mp, shape, tile):
    """Small-integer A and B keep every partial sum exact in FP64, so the
    result must equal tril(A) @ B bit for bit, whatever order the MFMA
    steps add in.  A carries nonzero values above the diagonal, which must
    not reach C.  Covers each tile shape, odd k starts (the boxes shape's
    triangles) and partial edge tiles; a wrong pairing of A and B values
    inside a k panel, or a mask off by one, changes integers."""
    plan = uvt.Plan(n, m, p, mp, shape)
    assert plan.tile_shape() == tile
    plan.close()
    rng = np.random.default_rng(n * 7919 + m)
    A = rng.integers(-4, 5, size=(n, n)).astype(np.float64)
    B = rng.integers(-4, 5, size=(n, m)).astype(np.float64)
    C = run_trmm(gpu, uvt, A, B, p, mp, shape)
    np.testing.assert_array_equal(C, np.tril(A) @ B)


CASES = [
    (2048, 2048, 4, "reg", "boxes"),  # the CPU-runnable reference configuration
    (2048, 2048, 4, "reg", "trapezoid"),
    (1000, 777, 3, "balanced", "boxes"),
    (513, 129, 16, "balanced", "trapezoid"),
    (127, 1, 2, "reg", "boxes"),
    (64, 300, 1, "reg", "trapezoid"),
    (4097, 160, 8, "balanced", "boxes"),
    (3000, 250, 7, "balanced", "trapezoid"),
]


@pytest.mark.parametrize("n,m,p,mp,shape", CASES)
def test_parity_seeded(gpu, uvt, oracle, n, m, p, mp, shape):
    A = oracle.seeded_A(n, 100 + n, 0)
    B = oracle.seeded_B(n, m, 100 + n, 0)
    C = run_trmm(gpu, uvt, A, B, p, mp, shape)
    err = rel_fro(C, oracle.trmm(A, B))
    say("rel_fro", err)
    assert err <= TOL


def test_parity_exponent_spread(gpu, uvt, oracle):
    n, m = 1500, 300
    A = oracle.seeded_A(n, 7, 1)
    B = oracle.seeded_B(n, m, 7, 1)
    C = run_trmm(gpu, uvt, A, B, 5, "balanced", "trapezoid")
    assert rel_fro(C, oracle.trmm(A, B)) <= TOL


def test_mkl_trmm_fixtures(gpu, uvt, oracle):
    import re
    d = np.load(os.path.join(GOLDEN, "mkl_trmm.npz"))
    for k in d.files:
        mt = re.fullmatch(r"s(\d+)_n(\d+)_m(\d+)_sp(\d+)_C", k)
        if not mt:
            continue
        seed, n, m, sp = map(int, mt.groups())
        A, B = oracle.seeded_A(n, seed, sp), oracle.seeded_B(n, m, seed, sp)
        for shape in ("boxes", "trapezoid"):
            C = run_trmm(gpu, uvt, A, B, min(4, n), "balanced" if n > 4 else "reg", shape)
            assert rel_fro(C, d[k]) <= TOL, (k, shape)


def test_mkl_stage_loop_fixtures(gpu, uvt, oracle):
    """The reference's own per-rank dgemm+dtrmm sequence (run with MKL) on its
    closed-form inputs, same band count and mapping, both shapes."""
    d = np.load(os.path.join(GOLDEN, "mkl_stage_loop.npz"))
    for k in d.files:
        if not k.endswith("_C"):
            continue
        n, p, mp, _ = k.split("_")
        n, p = int(n[1:]), int(p[1:])
        A, B = oracle.closed_A(n), oracle.closed_B(n, n)
        for shape in ("boxes", "trapezoid"):
            C = run_trmm(gpu, uvt, A, B, p, mp, shape)
            assert rel_fro(C, d[k]) <= TOL, (k, shape)


def test_generators_bit_exact(gpu, uvt, oracle):
    n, m = 333, 517
    dA, dB = gpu.alloc(n * n * 8), gpu.alloc(n * m * 8)
    say("fill closed form", n, m)
    gpu.fill_closed_form(n, m, dA, n, dB, m)
    gpu.sync()
    np.testing.assert_array_equal(dA.download((n, n)), oracle.closed_A(n))
    np.testing.assert_array_equal(dB.download((n, m)), oracle.closed_B(n, m))
    for spread in (0, 1):
        say("fill seeded", spread)
        gpu.fill_seeded(n, m, 42, spread, dA, n, dB, m)
        gpu.sync()
        np.testing.assert_array_equal(dA.download((n, n)), oracle.seeded_A(n, 42, spread))
        np.testing.assert_array_equal(dB.download((n, m)), oracle.seeded_B(n, m, 42, spread))
    dA.free()
    dB.free()


@pytest.mark.parametrize("n", [257, 1000, 2048, 4096, 8192, 16384, 32768])
def test_reference_norm_on_gpu(gpu, uvt, n):
    """Closed-form inputs generated on the GPU, product, norm: the reference's Norm line."""
    with open(os.path.join(GOLDEN, "reference_norms.json")) as f:
        want = {d["n"]: d["norm"] for d in json.load(f)["norms"]}[n]
    dA, dB, dC = gpu.alloc(n * n * 8), gpu.alloc(n * n * 8), gpu.alloc(n * n * 8)
    say("norm case", n)
    gpu.fill_closed_form(n, n, dA, n, dB, n)
    cases = (("reg", "boxes", 4), ("balanced", "trapezoid", 8)) if n <= 4096 else \
        (("balanced", "trapezoid", 16),)
    for mp, shape, p in cases:
        plan = uvt.Plan(n, n, p, mp, shape)
        gpu.trmm_f64(plan, dA, n, dB, n, dC, n)
        assert "%.3f" % gpu.inf_norm(n, n, dC, n) == want
        plan.close()
    for b in (dA, dB, dC):
        b.free()


def test_inf_norm_and_row_sums(gpu, uvt, oracle):
    C = oracle.seeded_B(1000, 3001, 5, 1)
    dC = gpu.array(C)
    v, rs = oracle.inf_norm(C)
    assert abs(gpu.inf_norm(1000, 3001, dC, 3001) - v) <= 1e-14 * v
    out = gpu.alloc(1000 * 8)
    gpu.row_abs_sums(1000, 3001, dC, 3001, out)
    gpu.sync()
    np.testing.assert_allclose(out.download(1000), rs, rtol=1e-14)
    dC.free()
    out.free()


@pytest.mark.parametrize("pad", [1, 2])
def test_odd_leading_dimension_and_upper_garbage(gpu, uvt, oracle, pad):
    """Junk above the diagonal (huge values, NaN) and in the row padding must
    never reach C: lda odd (unaligned rows, 8-byte loads) or even (16-byte
    loads, where a diagonal panel loads both elements of a pair and zeroes
    the one above the diagonal)."""
    n, m = 700, 90
    A = oracle.seeded_A(n, 3)
    B = oracle.seeded_B(n, m, 3)
    upper = np.triu(np.ones((n, n), dtype=bool), 1)
    Abig = np.full((n, n + pad), np.nan)
    Abig[:, :n] = np.where(upper, 1e300, A)
    Abig[:, :n][upper & (np.arange(n)[None, :] % 3 == 0)] = np.nan
    ref = oracle.trmm(A, B)
    for shape in ("boxes", "trapezoid"):
        C = run_trmm(gpu, uvt, Abig, B, 3, "reg", shape, lda=n + pad)
        assert rel_fro(C, ref) <= TOL


def test_boxes_equals_trapezoid_and_repeatable(gpu, uvt, oracle):
    n, m = 2500, 300
    A, B = oracle.seeded_A(n, 9), oracle.seeded_B(n, m, 9)
    c1 = run_trmm(gpu, uvt, A, B, 6, "balanced", "boxes")
    c2 = run_trmm(gpu, uvt, A, B, 6, "balanced", "trapezoid")
    c3 = run_trmm(gpu, uvt, A, B, 6, "balanced", "trapezoid")
    assert rel_fro(c1, c2) <= 1e-14
    np.testing.assert_array_equal(c2, c3)


@pytest.mark.parametrize("n,m,p,mapping,shape,pad,how", [
    # 128x128 tiles, few and deep (narrow B): stream-K
    (8165, 939, 16, "balanced", "trapezoid", 0, "sk"),  # 512 tiles, depths 128..8165
    (5000, 2000, 1, "reg", "trapezoid", 1, "sk"),       # odd lda (8-byte loads), an 8-row last row tile
    (6772, 1548, 4, "reg", "trapezoid", 0, "sk"),
    # 64x64 tiles, few and deep: whole tiles cut into k pieces
    (9000, 250, 1, "balanced", "trapezoid", 0, "cut"),
    (12000, 128, 4, "reg", "trapezoid", 0, "cut"),
    # whole tiles
    (4096, 4096, 1, "reg", "trapezoid", 0, "whole"),    # 128x128 tiles that balance by themselves
    (2048, 2048, 4, "reg", "boxes", 0, "whole"),        # the reference's N=2048 configuration
    (1030, 37, 1, "reg", "trapezoid", 0, "whole"),      # under 2e9 flops
])
def test_split_reductions(gpu, uvt, oracle, n, m, p, mapping, shape, pad, how):
    """Trapezoid plans whose whole tiles would not finish together split
    their reductions: with 128x128 tiles stream-K (up to one workgroup per
    slot, in groups that run one row tile's column tiles side by side, each
    the same number of k panels), with smaller tiles the deepest tiles
    cut into k pieces; the pieces after the first of a tile are added to C
    afterwards in k order.  Small-integer inputs give exact sums in any
    order, so they show every k panel of every tile counted exactly once;
    seeded inputs against exact double-double rows, and the same bits on
    every run."""
    plan = uvt.Plan(n, m, p, mapping, shape)
    wg, parts = gpu.plan_launch_info(plan)
    plan.close()
    say("launch", (n, m, p, mapping, shape), "workgroups", wg, "partial tiles", parts)
    slots = 2 * gpu.cus
    if how == "sk":
        # column groups of ceil(m / 128) workgroups, as many as fit the slots
        assert slots - (m + 127) // 128 < wg <= slots and parts > 0
    elif how == "cut":
        assert wg > slots and parts > 0
    else:
        assert parts == 0
    lda = n + pad
    rng = np.random.default_rng(n + m)
    Ai = np.tril(rng.integers(-3, 4, (n, n)).astype(np.float64))
    Bi = rng.integers(-3, 4, (n, m)).astype(np.float64)
    Ap = np.zeros((n, lda))
    Ap[:, :n] = Ai
    np.testing.assert_array_equal(run_trmm(gpu, uvt, Ap, Bi, p, mapping, shape, lda=lda), Ai @ Bi)
    del Ai, Bi, Ap
    A, B = oracle.seeded_A(n, 31, 1), oracle.seeded_B(n, m, 31, 1)
    c1 = run_trmm(gpu, uvt, A, B, p, mapping, shape)
    c2 = run_trmm(gpu, uvt, A, B, p, mapping, shape)
    np.testing.assert_array_equal(c1, c2)
    rows = np.unique(np.concatenate([[0, 1, 63, 64, 127, 128, n // 2, n - 2, n - 1],
                                     np.arange(0, n, 512), rng.integers(0, n, 24)]))
    err = rel_fro(c1[rows], oracle.trmm_rows_dd(A, B, rows))
    say("split reductions", how, (n, m, p, mapping, shape), "err", err)
    assert err <= 1e-13
    if shape == "boxes":
        c3 = run_trmm(gpu, uvt, A, B, p, mapping, "trapezoid")
        assert rel_fro(c1, c3) <= 1e-14


def test_headline_plan_keeps_whole_tiles(gpu, uvt):
    plan = uvt.Plan(32768, 8192, 16, "balanced", "trapezoid")
    wg, parts = gpu.plan_launch_info(plan)
    plan.close()
    assert (wg, parts) == (128 * 64, 0)  # 256x128 tiles, whole


def test_large_row_sampled(gpu, uvt, oracle):
    """N=16384 balanced/trapezoid over 16 bands (the single-GPU reference
    configuration's shape) with B cut to 512 columns; rows checked against
    double-double sums."""
    n, m = 16384, 512
    dA, dB, dC = gpu.alloc(n * n * 8), gpu.alloc(n * m * 8), gpu.alloc(n * m * 8)
    say("large case", n, m)
    gpu.fill_seeded(n, m, 77, 0, dA, n, dB, m)
    plan = uvt.Plan(n, m, 16, "balanced", "trapezoid")
    gpu.trmm_f64(plan, dA, n, dB, m, dC, m)
    gpu.sync()
    rows = np.array([0, 1, 4094, 4095, 4096, 8191, 12000, 15862, 15863, 16383])
    A = oracle.seeded_A(n, 77)
    B = oracle.seeded_B(n, m, 77)
    ref = oracle.trmm_rows_dd(A, B, rows)
    C = dC.download((n, m))
    for r, i in enumerate(rows):
        assert rel_fro(C[i], ref[r]) <= TOL, i
    # the rest through a size-independent property: linearity in B
    dB2, dC2 = gpu.alloc(n * m * 8), gpu.alloc(n * m * 8)
    gpu.fill_seeded(n, m, 78, 0, None, n, dB2, m)
    B2 = dB2.download((n, m))
    dB.upload(B + B2)
    gpu.trmm_f64(plan, dA, n, dB, m, dC2, m)
    gpu.trmm_f64(plan, dA, n, dB2, m, dB, m)
    gpu.sync()
    lhs, rhs = dC2.download((n, m)), C + dB.download((n, m))
    assert rel_fro(lhs, rhs) <= 1e-13
    plan.close()
    for b in (dA, dB, dC, dB2, dC2):
        b.free()


def test_cli_reference_output(uvt):
    """uvatrmm.x 2048 reg boxes with 4 bands prints the reference's lines."""
    say("cli 2048 reg boxes --bands 4")
    import tempfile
    cwd = tempfile.mkdtemp()
    out = subprocess.run([uvt.CLI_PATH, "2048", "reg", "boxes", "--bands", "4"], capture_output=True,
                         text=True, timeout=120, cwd=cwd)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.split("\n")
    assert lines[0] == "" and lines[1].startswith("Total t: ")
    assert lines[2] == "Total mpis: 4" and lines[3].startswith("Mult t: ")
    assert "\nNorm: 61.447\n" in out.stdout
    with open(os.path.join(cwd, "norm_ff.txt")) as f:
        assert f.read() == "61.447"
    out = subprocess.run([uvt.CLI_PATH, "1000", "balanced", "trapezoid", "--bands", "3"],
                         capture_output=True, text=True, timeout=120, cwd=cwd)
    assert "\nNorm: 29.286\n" in out.stdout


def test_cli_result_dump(uvt, oracle):
    """--dump writes C in the reference's per-rank binary layout: header
    {rows, columns, 0, rows, column begin, column count}, then the rank's
    column slice row by row, the slices being map_regular over the columns."""
    import tempfile
    n, m, p = 300, 100, 3
    say("cli dump", n, m, p)
    cwd = tempfile.mkdtemp()
    out = subprocess.run([uvt.CLI_PATH, str(n), "reg", "boxes", "--bands", str(p), "--m", str(m),
                          "--dump", cwd], capture_output=True, text=True, timeout=120, cwd=cwd)
    assert out.returncode == 0, out.stderr
    ref = oracle.trmm(oracle.closed_A(n), oracle.closed_B(n, m))
    begin = 0
    for r in range(p):
        raw = open(os.path.join(cwd, f"output_uvatrmm.{n}.{p}.{r}.dat"), "rb").read()
        hdr = np.frombuffer(raw[:24], dtype=np.int32)
        size = m // p + (m % p if r == p - 1 else 0)
        assert hdr.tolist() == [n, m, 0, n, begin, size]
        data = np.frombuffer(raw[24:], dtype=np.float64).reshape(n, size)
        assert rel_fro(data, ref[:, begin:begin + size]) <= 1e-14
        begin += size
    assert begin == m


def test_cli_split_precision_norm(uvt):
    """uvatrmm.x with --prec split (default 13 moduli) on the reference's
    closed-form inputs prints the reference's Norm line (the split path's
    ~1e-13 relative error is far below the printed 3 decimals)."""
    import tempfile
    cwd = tempfile.mkdtemp()
    for args, want in ((["2048", "reg", "boxes", "--bands", "4"], "61.447"),
                       (["4096", "balanced", "trapezoid", "--bands", "8", "--slices", "12"], "243.282")):
        say("cli split", args)
        out = subprocess.run([uvt.CLI_PATH, *args, "--prec", "split"], capture_output=True,
                             text=True, timeout=120, cwd=cwd)
        assert out.returncode == 0, out.stderr
        assert f"\nNorm: {want}\n" in out.stdout, out.stdout


def test_max_order_65536(gpu, uvt, oracle):
    """N = 65536 (A alone is 34 GB; row offsets pass 2^32 elements), balanced
    over 16 bands, trapezoid, M = 64: sampled rows against double-double sums
    of the seeded inputs generated on the host element by element."""
    n, m = 65536, 64
    say("max order", n, m)
    dA, dB, dC = gpu.alloc(n * n * 8), gpu.alloc(n * m * 8), gpu.alloc(n * m * 8)
    try:
        gpu.fill_seeded(n, m, 91, 1, dA, n, dB, m)
        plan = uvt.Plan(n, m, 16, "balanced", "trapezoid")
        gpu.trmm_f64(plan, dA, n, dB, m, dC, m)
        gpu.sync()
        bands = plan.bands()
        plan.close()
        C = dC.download((n, m))
    finally:
        for b in (dA, dB, dC):
            b.free()
    rows = sorted({0, 1, 32767, 32768, 65535} | {b for b, _ in bands[1:4]} | {b - 1 for b, _ in bands[1:4]})
    ref = oracle.seeded_rows_dd(n, m, 91, 1, rows)
    for r, i in enumerate(rows):
        assert rel_fro(C[i], ref[r]) <= TOL, i


def test_random_plans_exact(gpu, uvt):
    """A seeded sweep of plans: orders log-uniform in [1, 4500], B widths in
    [1, 2500], both mappings and shapes, then deep narrow trapezoids (n in
    [5000, 9000], m in [100, 1300]); 1..16 bands, odd and even leading
    dimensions.  Small-integer inputs make every product exact in any
    summation order, so each result must equal NumPy's bit for bit, whatever
    the plan launches on this GPU (whole tiles, cut reductions, stream-K,
    boxes).  Plans the mapping refuses (a band without rows) are skipped."""
    rng = np.random.default_rng(2026)
    seen = set()
    for case in range(42):
        if case < 36:
            n = int(np.exp(rng.uniform(0, np.log(4500))))
            m = int(np.exp(rng.uniform(0, np.log(2500))))
            shape = ("trapezoid", "boxes")[(case // 2) % 2]
        else:  # deep and narrow trapezoids: their reductions are split
            n = int(rng.integers(5000, 9001))
            m = int(rng.integers(100, 1301))
            shape = "trapezoid"
        p = int(rng.integers(1, min(16, n) + 1))
        mapping = ("reg", "balanced")[case % 2]
        pad = int(rng.integers(0, 2))
        try:
            plan = uvt.Plan(n, m, p, mapping, shape)
        except uvt.UvtError as e:
            assert e.status == uvt.ERR_MAPPING
            continue
        wg, parts = gpu.plan_launch_info(plan)
        plan.close()
        how = "whole" if parts == 0 else ("stream-k" if wg == 2 * gpu.cus and shape == "trapezoid" else "cut")
        seen.add(how)
        say("random plan", case, (n, m, p, mapping, shape, pad), how, wg, parts)
        A = np.tril(rng.integers(-2, 3, (n, n)).astype(np.float64))
        B = rng.integers(-2, 3, (n, m)).astype(np.float64)
        Ap = np.zeros((n, n + pad))
        Ap[:, :n] = A
        C = run_trmm(gpu, uvt, Ap, B, p, mapping, shape, lda=n + pad)
        np.testing.assert_array_equal(C, A @ B, err_msg=str((n, m, p, mapping, shape, pad)))
    say("launch forms seen", sorted(seen))
    assert "whole" in seen and len(seen) >= 2
