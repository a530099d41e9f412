"""The reconstruction's arithmetic (k_reconstruct in uvatrmm_amd/csrc/uvt_split.hip),
modelled step by step with exactly rounded FP64 operations and checked against
the exact Chinese remainder value from Python integers, for every moduli count
the library accepts (2..14):

* every limb sum, quotient and carry it forms is an integer below 2^53, so the
  FP64 steps before the last two are exact; with X_0 brought into [0, 2^40),
  X_2 2^40 + X_1 is the integer floor(C' / 2^40) whatever the split between
  X_1 and X_2, so the last two steps round it and then C' once each;
* the result is within one ulp of C', and equal to it whenever C' is an FP64
  number (the integer-input exactness the GPU tests rely on);
* this holds up to |C'| = M/2 (1 - 1.4e-6), the bound the scale target keeps.

The GPU tests check the kernel itself against the FP64 oracle; this pins the
method's error claims, which those tolerances cannot see."""
import math
import random
from fractions import Fraction

# pairwise coprime moduli, largest first (the library's list)
MODS = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199]
TWO40 = 1099511627776.0
INV40 = 1.0 / TWO40
INV80 = INV40 * INV40


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))  # one correct rounding


def tables(n):
    M = math.prod(MODS[:n])
    mask = (1 << 40) - 1
    e = [[0.0] * n for _ in range(3)]
    for l in range(n):
        ml = MODS[l]
        W = M // ml
        inv = pow(W % ml, -1, ml)
        el = W * inv
        e[0][l], e[1][l], e[2][l] = float(el & mask), float((el >> 40) & mask), float(el >> 80)
    mk = [float(M & mask), float((M >> 40) & mask), float(M >> 80)]
    inv_mhi = 1.0 / (mk[2] + mk[1] / TWO40 + mk[0] / (TWO40 * TWO40))
    return M, e, mk, inv_mhi


def reconstruct(res, tab):
    _, e, mk, inv_mhi = tab
    s0 = s1 = s2 = 0.0
    for l, r in enumerate(res):
        s0, s1, s2 = fma(r, e[0][l], s0), fma(r, e[1][l], s1), fma(r, e[2][l], s2)
    q = float(round(fma(s1, INV40, fma(s0, INV80, s2)) * inv_mhi))
    x0, x1, x2 = fma(-q, mk[0], s0), fma(-q, mk[1], s1), fma(-q, mk[2], s2)
    c0 = math.floor(x0 * INV40)
    x0, x1 = fma(-c0, TWO40, x0), x1 + c0
    steps = (s0, s1, s2, q, x0, x1, x2, c0)
    return fma(fma(x2, TWO40, x1), TWO40, x0), steps


def test_limb_crt_matches_exact_crt():
    rng = random.Random(7)
    for n in range(2, 15):
        tab = tables(n)
        M = tab[0]
        lim = int(Fraction(M, 2) * (1 - Fraction(14, 10**7)))
        cases = [0, 1, -1, lim, -lim, lim - 1, -lim + 1] + [rng.randint(-lim, lim) for _ in range(60)]
        for k in range(0, M.bit_length() - 2, 3):
            for s in (1, 3, 12345, (1 << 52) + 1):
                if (s << k) <= lim:
                    cases += [s << k, -(s << k)]
        for x in cases:
            got, steps = reconstruct([x % MODS[l] for l in range(n)], tab)
            for v in steps:
                assert v == math.floor(v) and abs(v) < 2**53, (n, x, steps)
            ref = float(x)
            assert abs(Fraction(got) - x) <= Fraction(math.ulp(ref)), (n, x, got)
            if Fraction(ref) == x:
                assert got == ref, (n, x, got)
