"""The serving loop's host logic without a GPU: the arena's first-fit blocks
and the FIFO admission rule driven by a fake executor (tests/serve_arena_check.cpp,
compiled here with the host compiler against uvatrmm_amd/csrc/uvt_serve_arena.hpp)."""
import os
import shutil
import subprocess
import tempfile

import pytest

from conftest import ROOT


def test_arena_and_admission():
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    out = os.path.join(tempfile.mkdtemp(), "serve_arena_check")
    subprocess.check_call([cxx, "-std=c++17", "-O1", "-Wall", "-Werror",
                           "-I", os.path.join(ROOT, "uvatrmm_amd", "csrc"),
                           os.path.join(ROOT, "tests", "serve_arena_check.cpp"), "-o", out])
    r = subprocess.run([out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "ok"
