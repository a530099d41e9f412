"""Host-side sizing of the split path (no GPU): the workspace of the split
headline (N = 32768, M = 8192, 12 moduli) and its scaling."""
import ctypes

import pytest


def workspace(uvt, n, m, s):
    b = ctypes.c_size_t()
    st = uvt.lib().uvt_trmm_split_workspace(n, m, s, ctypes.byref(b))
    if st != uvt.OK:
        raise uvt.UvtError(st)
    return b.value


def test_headline_workspace_below_11_5_gib(uvt):
    """A' residues are stored only over each 256-row tile's box and B'
    residues in two groups of moduli: 12 moduli at N = 32768, M = 8192 need
    the residues of A (6.4 GB), half of B's (1.6 GB), the product residues
    (3.2 GB) and the scales and error terms."""
    gib = workspace(uvt, 32768, 8192, 12) / 2**30
    assert 10.0 <= gib <= 11.5, gib


def test_workspace_grows_with_moduli_and_columns(uvt):
    assert workspace(uvt, 8192, 1024, 13) > workspace(uvt, 8192, 1024, 12)
    assert workspace(uvt, 8192, 2048, 12) > workspace(uvt, 8192, 1024, 12)
    assert workspace(uvt, 200, 50, 12) > 0  # no box: only the error-statement terms


def test_workspace_refusals(uvt):
    with pytest.raises(uvt.UvtError):
        workspace(uvt, 131073, 8, 14)
    with pytest.raises(uvt.UvtError):
        workspace(uvt, 65536, 8, 2)  # too few moduli for that depth
