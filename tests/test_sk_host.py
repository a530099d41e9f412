"""The stream-K geometry of the trapezoid walk without a GPU: piece coverage,
partial-tile slots, the fixup rule and the XCD mapping of the kernels
(tests/sk_geometry_check.cpp, compiled here with the host compiler against
uvatrmm_amd/csrc/uvt_sk.hpp)."""
import os
import shutil
import subprocess
import tempfile

import pytest

from conftest import ROOT


def test_stream_k_geometry():
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    out = os.path.join(tempfile.mkdtemp(), "sk_geometry_check")
    subprocess.check_call([cxx, "-std=c++17", "-O2", "-Wall", "-Werror",
                           "-I", os.path.join(ROOT, "uvatrmm_amd", "csrc"),
                           os.path.join(ROOT, "tests", "sk_geometry_check.cpp"), "-o", out])
    r = subprocess.run([out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "ok"
