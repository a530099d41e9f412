"""Host-side planner of libuvt (no GPU): bit-exact mapping and tile plans."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _golden():
    with open(os.path.join(GOLDEN, "mapping_golden.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("case", _golden()["bands"], ids=lambda c: f"{c['n']}-{c['p']}-{c['map']}")
def test_product_mapping_golden(uvt, case):
    want = [tuple(int(x) for x in b.split(":")) for b in case["bands"].split(",")]
    assert uvt.plan_bands(case["n"], case["p"], case["map"]) == want


def test_product_mapping_equals_oracle(uvt, oracle):
    rng = np.random.default_rng(1)
    ns = list(range(1, 70)) + [int(x) for x in rng.integers(70, 92000, 300)]
    for n in ns:
        for p in (1, 2, 3, 4, 5, 8, 13, 16, 31, 64):
            if p > n:
                continue
            for mp in ("reg", "balanced"):
                rc, ref = oracle.bands(n, p, mp)
                if rc == 0:
                    assert uvt.plan_bands(n, p, mp) == ref, (n, p, mp)
                else:
                    with pytest.raises(uvt.UvtError) as e:
                        uvt.plan_bands(n, p, mp)
                    assert e.value.ref_exit == rc, (n, p, mp)


def test_mapping_errors(uvt):
    with pytest.raises(uvt.UvtError) as e:
        uvt.plan_bands(7, 7, "balanced")
    assert e.value.status == uvt.ERR_MAPPING and e.value.ref_exit == 11
    with pytest.raises(uvt.UvtError) as e:
        uvt.plan_bands(3, 4, "reg")
    assert e.value.status == uvt.ERR_MAPPING and e.value.ref_exit == 3
    with pytest.raises(uvt.UvtError) as e:
        uvt.plan_bands(92682, 4, "balanced")  # 32-bit element count would overflow
    assert e.value.status == uvt.ERR_SIZE
    assert uvt.plan_bands(92681, 4, "balanced")[0][0] == 0


def _coverage(plan, n, m):
    """For every C element: how many tiles write it (store) and the k set they sum."""
    tm, tn = plan.tile_shape()
    store = np.zeros((n, m), np.int32)
    kmin = np.full((n, m), 1 << 30, np.int64)
    kmax = np.full((n, m), -1, np.int64)
    ksum = np.zeros((n, m), np.int64)
    for ph in range(plan.phases):
        for t in plan.tiles(ph):
            r0, r1 = t["r0"], t["r0"] + t["rows"]
            c0, c1 = t["c0"], min(t["c0"] + tn, m)
            assert 0 < t["rows"] <= tm and 0 <= r0 and r1 <= n and c0 < m
            for r in range(r0, r1):
                hi = min(t["k_hi"], r + 1) if t["flags"] & uvt.TILE_LOWER else t["k_hi"]
                lo = t["k_lo"]
                assert hi <= r + 1, "a tile reads A above the diagonal"
                assert lo < hi, "zero tile emitted"
                if not (t["flags"] & uvt.TILE_ACCUM):
                    store[r, c0:c1] += 1
                kmin[r, c0:c1] = np.minimum(kmin[r, c0:c1], lo)
                kmax[r, c0:c1] = np.maximum(kmax[r, c0:c1], hi)
                ksum[r, c0:c1] += hi - lo
    return store, kmin, kmax, ksum


import uvatrmm_amd as uvt  # noqa: E402


@pytest.mark.parametrize("n,m,p,mp,shape", [
    (300, 200, 4, "reg", "boxes"), (300, 200, 4, "balanced", "boxes"),
    (300, 200, 5, "balanced", "trapezoid"), (129, 257, 3, "reg", "trapezoid"),
    (1, 1, 1, "reg", "boxes"), (1, 5, 1, "reg", "trapezoid"), (10, 3, 4, "balanced", "boxes"),
    (500, 64, 16, "balanced", "boxes"),
])
def test_plan_covers_lower_trapezoid_exactly(uvt, n, m, p, mp, shape):
    """Every C element is stored once and sums exactly k in [0, row] once."""
    plan = uvt.Plan(n, m, p, mp, shape)
    store, kmin, kmax, ksum = _coverage(plan, n, m)
    rows = np.arange(n)[:, None]
    assert np.all(store == 1)
    assert np.all(kmin == 0)
    assert np.all(kmax == rows + 1)
    assert np.all(ksum == rows + 1)  # disjoint k ranges covering [0, row]


def test_boxes_geometry_follows_the_reference(uvt):
    """boxes: per band a GEMM box with K = b+1 then the (h-1) triangle
    (UVaTRMM/src/mm_triangular.c:347-387)."""
    n, m, p = 2048, 512, 4
    plan = uvt.Plan(n, m, p, "reg", "boxes")
    bands = plan.bands()
    assert bands == [(0, 512), (512, 512), (1024, 512), (1536, 512)]
    rect, tri = plan.tiles(0), plan.tiles(1)
    for s, (b, h) in enumerate(bands):
        r = rect[rect["band"] == s]
        assert set(r["k_hi"]) == {b + 1} and set(r["k_lo"]) == {0}
        tn = plan.tile_shape()[1]
        assert r["rows"].sum() == h * ((m + tn - 1) // tn)
        t = tri[tri["band"] == s]
        assert set(t["k_lo"]) == {b + 1}
        assert np.all(t["flags"] == uvt.TILE_LOWER | uvt.TILE_ACCUM)
        assert t["r0"].min() == b + 1 and (t["r0"] + t["rows"]).max() == b + h


def test_tiles_heavy_first(uvt):
    plan = uvt.Plan(4096, 1024, 16, "balanced", "trapezoid")
    t = plan.tiles(0)
    work = t["k_hi"] - t["k_lo"]
    assert np.all(np.diff(work) <= 0)
    assert plan.issued_flops >= plan.flops
    assert plan.flops == 1024.0 * 4096 * 4096
    # masked diagonal waste stays small at this size (tile 128 vs n 4096)
    assert plan.issued_flops / plan.flops < 1.05


def test_plan_argument_errors(uvt):
    for args in ((0, 4, 1), (4, 0, 1), (4, 4, 0)):
        with pytest.raises(uvt.UvtError) as e:
            uvt.Plan(*args)
        assert e.value.status == uvt.ERR_ARG
