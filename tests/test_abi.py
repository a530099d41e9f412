"""The C ABI without a GPU: the library loads, exports every declared entry
point, and the command line keeps the reference's argument errors."""
import ctypes
import os
import re
import subprocess
import sys
import tempfile

import pytest

from conftest import ROOT

HEADERS = ["uvt.h", "uvt_cblas.h"]


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(uvt_\w+|cblas_\w+)\s*\(", text)
    return sorted(set(n for n in names if not n.endswith("_s")))


@pytest.mark.parametrize("header", HEADERS)
def test_library_exports_declared_symbols(uvt, header):
    lib = ctypes.CDLL(uvt.LIB_PATH if header != "uvt_cblas.h" else uvt.CBLAS_PATH)
    names = declared(header)
    assert len(names) > 10 or header != "uvt.h"
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("header", HEADERS)
def test_library_exports_nothing_undeclared(uvt, header):
    """The exported entry points are exactly the header's: library-internal
    helpers stay hidden."""
    path = uvt.LIB_PATH if header != "uvt_cblas.h" else uvt.CBLAS_PATH
    try:
        out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                             timeout=60, check=True).stdout
    except (OSError, subprocess.CalledProcessError):
        pytest.skip("nm not available")
    exported = set()
    for line in out.splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[1] == "T" and re.match(r"(uvt_|cblas_)", parts[2]):
            exported.add(parts[2])
    extra = sorted(exported - set(declared(header)))
    assert not extra, extra
    # C++ symbols of the library's own (kernels, their launch stubs, helper
    # functions and types) are hidden too; only instantiations of the
    # standard library's templates may remain
    cxx = subprocess.run(["nm", "-DC", "--defined-only", path], capture_output=True, text=True,
                         timeout=60, check=True).stdout
    own = []
    for line in cxx.splitlines():
        parts = line.split(None, 2)
        if len(parts) < 3 or parts[1] not in "TtWVDBR":
            continue
        name = parts[2]
        if name.startswith(("std::", "vtable for std::", "typeinfo for std::",
                            "typeinfo name for std::", "void std::", "__gnu_cxx::", "__hip_cuid_")):
            continue
        if "(" in name or "::" in name or name.startswith(("k_", "__device_stub")):
            own.append(name)
    assert not own, own


def test_version_and_status_strings(uvt):
    assert b"gfx950" in uvt.lib().uvt_version()
    assert uvt.lib().uvt_status_string(uvt.ERR_MAPPING).startswith(b"mapping")


def _cli(*args):
    return subprocess.run([os.path.join(ROOT, "uvatrmm_amd", "bin", "uvatrmm.x"), *args],
                          capture_output=True, text=True, timeout=60, cwd=tempfile.mkdtemp())


def test_cli_usage_errors_exit_3(uvt):
    for args, msg in ((["2048", "reg"], "Number of arguments"),
                      (["2048", "foo", "boxes"], "Unknown name of mapping function"),
                      (["2048", "reg", "full"], "Unknown name of communication shape"),
                      (["3", "reg", "boxes", "--bands", "4"],
                       "Number of rows should be equal or greater than number of processors")):
        r = _cli(*args)
        assert r.returncode == 3, (args, r.returncode)
        assert f"Error: {msg}" in r.stderr
        assert "Usage:" in r.stderr


def test_cli_mapping_abort_codes(uvt):
    assert _cli("7", "balanced", "boxes", "--bands", "7").returncode == 11


def test_split_routing_falls_back_to_fp64(uvt):
    """A split-class request too deep for its moduli (n = 65536 with 2 moduli
    leaves fewer than 6 bits per operand) runs in FP64 instead of failing:
    the server charges it no split workspace."""
    def nbytes(n, m, prec, slices):
        import ctypes
        cfg = uvt.ServeConfig(1 << 40, 1, slices, None, 0.0, 0, 0)
        b = ctypes.c_size_t()
        uvt.check(uvt.lib().uvt_serve_request_bytes(ctypes.byref(cfg), n, m, prec, ctypes.byref(b)))
        return b.value
    assert nbytes(65536, 2048, uvt.PREC_SPLIT, 2) == nbytes(65536, 2048, uvt.PREC_F64, 2)
    assert nbytes(65536, 2048, uvt.PREC_SPLIT, 13) > nbytes(65536, 2048, uvt.PREC_F64, 13)


def test_serve_request_bytes(uvt):
    """Arena bytes of a request: A, B and C (256-byte aligned pieces), plus the
    split workspace when a split-class request takes the int8 path (where it
    is faster on its own: m >= 1024 and n >= 4096, or m >= 512 and
    n >= 6144); otherwise it runs in FP64 and needs none."""
    a = lambda x: (x + 255) // 256 * 256
    for n, m in [(100, 7), (768, 63), (2000, 500)]:
        assert uvt.serve_request_bytes(n, m, uvt.PREC_F64) == a(n * n * 8) + 2 * a(n * m * 8)
    for n, m in [(2000, 500), (8192, 128), (3072, 4096), (4096, 512), (6144, 256)]:  # FP64 is faster
        assert uvt.serve_request_bytes(n, m, uvt.PREC_SPLIT) == \
            uvt.serve_request_bytes(n, m, uvt.PREC_F64)
    for n, m in [(4096, 1024), (6144, 512), (8192, 4096)]:  # int8 is faster
        ws = ctypes.c_size_t()
        assert uvt.lib().uvt_trmm_split_workspace(n, m, 13, ctypes.byref(ws)) == uvt.OK
        assert uvt.serve_request_bytes(n, m, uvt.PREC_SPLIT) == \
            uvt.serve_request_bytes(n, m, uvt.PREC_F64) + a(ws.value)
    with pytest.raises(uvt.UvtError):
        uvt.serve_request_bytes(0, 5, uvt.PREC_F64)


def _cblas_subprocess(code):
    """Runs a ctypes snippet against libuvt_cblas in a child process."""
    prog = ("import ctypes, numpy as np\n"
            f"L = ctypes.CDLL({os.path.join(ROOT, 'uvatrmm_amd', 'lib', 'libuvt_cblas.so')!r})\n"
            "d, i = ctypes.c_double, ctypes.c_int\n" + code)
    return subprocess.run([sys.executable, "-c", prog], capture_output=True, text=True, timeout=120)


def test_cblas_illegal_parameters_return(uvt):
    """An illegal parameter is reported by position (CBLAS numbering) and the
    call returns with nothing written and no GPU touched, as MKL does."""
    code = """
C = np.full(16, 7.0); A = np.ones(16); B = np.ones(16)
p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
L.cblas_dgemm(101, 111, 111, -1, 2, 2, d(1), p(A), 2, p(B), 2, d(0), p(C), 2)
L.cblas_dgemm(999, 111, 111, 2, 2, 2, d(1), p(A), 2, p(B), 2, d(0), p(C), 2)
L.cblas_dgemm(101, 111, 111, 2, 3, 2, d(1), p(A), 2, p(B), 2, d(0), p(C), 3)
L.cblas_dgemm(102, 112, 111, 2, 2, 3, d(1), p(A), 2, p(B), 3, d(0), p(C), 2)
L.cblas_dtrmm(101, 141, 122, 111, 131, 4, 2, d(1), p(A), 3, p(B), 2)
L.cblas_dtrmm(101, 141, 122, 111, 133, 2, 2, d(1), p(A), 2, p(B), 2)
L.cblas_dtrmm(102, 142, 121, 111, 131, 3, 2, d(1), p(A), 2, p(B), 2)
assert (C == 7.0).all() and (B == 1.0).all()
print("ok")
"""
    r = _cblas_subprocess(code)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "ok"
    lines = [l for l in r.stderr.splitlines() if "libuvt_cblas" in l]
    assert [l.split("parameter ")[1].split()[0] for l in lines] == \
        ["4", "1", "11", "9", "10", "5", "12"], lines
    assert all("cblas_dgemm" in l for l in lines[:4]) and all("cblas_dtrmm" in l for l in lines[4:])


def test_split_refuses_too_few_moduli(uvt):
    """Two moduli cannot carry a reduction of 130000 (the scaled operands
    would keep no bits): the workspace query refuses the size instead of
    letting the product come out wrong.  Host-only check."""
    b = ctypes.c_size_t()
    assert uvt.lib().uvt_trmm_split_workspace(130000, 64, 2, ctypes.byref(b)) == uvt.ERR_SIZE
    assert uvt.lib().uvt_trmm_split_workspace(2000, 64, 2, ctypes.byref(b)) == uvt.OK
    assert uvt.lib().uvt_trmm_split_workspace(130000, 64, 12, ctypes.byref(b)) == uvt.OK
