"""Tile-shape choice (built-in rule and autotuned tables), host side only."""
import json
import os

import pytest


def test_builtin_choices(uvt):
    assert uvt.tune_choice(32768, 8192) == "128x128"
    assert uvt.tune_choice(4096, 16) == "128x32"
    assert uvt.tune_choice(300, 500) == "64x64"
    assert uvt.tune_choice(4096, 64) == "64x64"
    # alone, fewer than 512 tiles of 128x128 would leave CUs half empty
    assert uvt.tune_choice(2048, 2048) == "64x64"
    assert uvt.tune_choice(16384, 16384) == "128x128"


def test_table_lookup(uvt, tmp_path):
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"version": 1, "table": [
        {"n_max": 128, "m_max": 1024, "cfg": "64x64", "ms": {"128x128": 1.0}},
        {"n_max": 4096, "m_max": 64, "cfg": "128x32"},
        {"n_max": 4096, "m_max": 1024, "cfg": "128x128"}]}))
    assert uvt.tune_choice(100, 1000, str(p)) == "64x64"
    assert uvt.tune_choice(2000, 50, str(p)) == "128x32"
    assert uvt.tune_choice(2000, 500, str(p)) == "128x128"
    assert uvt.tune_choice(9000, 500, str(p)) == uvt.tune_choice(9000, 500)  # outside: built-in


def test_bad_table_rejected(uvt, tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{"table": [{"n_max": 1, "m_max": 2, "cfg": "99x99"}]}')
    with pytest.raises(uvt.UvtError):
        uvt.tune_choice(10, 10, str(p))
    with pytest.raises(uvt.UvtError):
        uvt.tune_choice(10, 10, str(tmp_path / "missing.json"))


def test_shipped_table_parses(uvt):
    path = os.path.join(os.path.dirname(uvt.__file__), "tune", "f64_tiles.json")
    if not os.path.exists(path):
        pytest.skip("no autotuned table shipped yet")
    for n in (64, 100, 1000, 4096):
        for m in (16, 100, 1024):
            assert uvt.tune_choice(n, m, path) in ("128x128", "64x64", "128x32")
