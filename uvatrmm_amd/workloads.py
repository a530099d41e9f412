"""Seeded synthetic workloads of the benchmark, shared with the tests so that
both run exactly the same problems.

batch_problems: 512 uneven TRMMs, orders log-uniform in [64, 4096], B
columns uniform in [16, 1024], a fixed seeded half of them split class.
serve_sizes: the served stream's requests, orders log-uniform in
[128, 16384], B columns uniform in [64, 4096], half split class.
"""
import numpy as np

PREC_F64, PREC_SPLIT = 0, 1


def batch_problems(seed, count=512, nmin=64, nmax=4096, mmin=16, mmax=1024):
    """[(n, m, prec)] of the batched workload; seed is the bench's --seed."""
    rng = np.random.default_rng(seed + 17)
    ns = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), count))).astype(int)
    ms = rng.integers(mmin, mmax + 1, count)
    order = rng.permutation(count)  # which half is split: a fixed seeded permutation
    split = set(order[: count // 2].tolist())
    return [(int(ns[k]), int(ms[k]), PREC_SPLIT if k in split else PREC_F64) for k in range(count)]


def serve_sizes(rng, count, nmin=128, nmax=16384, mmin=64, mmax=4096):
    """[(n, m, prec)] drawn from a numpy Generator (the stream's own)."""
    ns = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), count))).astype(int)
    ms = rng.integers(mmin, mmax + 1, count)
    precs = np.where(rng.random(count) < 0.5, PREC_SPLIT, PREC_F64)
    return [(int(a), int(b), int(c)) for a, b, c in zip(ns, ms, precs)]
