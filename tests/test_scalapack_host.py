"""The ScaLAPACK pdtrmm comparator (oracle/scalapack_trmm.c, a timing harness
around the call of UVaTRMM ScaLAPACK/scalapack_sample_trmm.c:135-139): on a
2 x 2 process grid with blocks smaller than the matrix, the norm it prints
equals the norm of the oracle's product of the same seeded inputs, also for a
column sample of a wider B (the generator indexes B by the full width)."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "..", "oracle", "scalapack_trmm.x")
MPIEXEC = "/opt/conda/bin/mpiexec"

pytestmark = pytest.mark.skipif(not (os.path.exists(EXE) and os.path.exists(MPIEXEC)),
                                reason="ScaLAPACK comparator or MPICH not built here")


def run(n, m, nb, seed, spread, mgen=None):
    cmd = [MPIEXEC, "-n", "4", EXE, str(n), str(m), str(nb), "2", "2", "1", str(seed), str(spread)]
    if mgen is not None:
        cmd.append(str(mgen))
    env = dict(os.environ, OMP_NUM_THREADS="1", MKL_NUM_THREADS="1")
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr[-400:]
    f = r.stdout.split()
    assert [int(x) for x in f[:6]] == [n, m, nb, 2, 2, 4]
    return float(f[9])


@pytest.mark.parametrize("n,m,nb,mgen", [(300, 70, 64, None), (257, 33, 32, 90), (130, 130, 256, None)])
def test_pdtrmm_norm_equals_oracle(oracle, n, m, nb, mgen):
    seed = 11
    fro = run(n, m, nb, seed, 1, mgen)
    A = oracle.seeded_A(n, seed, 1)
    B = oracle.seeded_B(n, mgen or m, seed, 1)[:, :m]
    ref = float(np.linalg.norm(oracle.trmm(A, np.ascontiguousarray(B))))
    assert abs(fro - ref) <= 1e-13 * ref, (fro, ref)


def test_rejects_a_grid_that_does_not_match_the_ranks():
    env = dict(os.environ, OMP_NUM_THREADS="1", MKL_NUM_THREADS="1")
    r = subprocess.run([MPIEXEC, "-n", "2", EXE, "64", "8", "16", "2", "2", "1", "1", "0"],
                       capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 2
