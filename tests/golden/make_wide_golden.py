"""Write r2_wide.npz: r^2 of get_pairwise_ld_score (gemm_cpu.c:343-404 of the
reference, compiled where it lies by ``make -C oracle ref``) for sample
counts N >= 2^24, where the reference's float casts of the counts and of N
round (it takes any int snp_size).  For each N: 24 column and 24 row bit
counts (random, plus 0, 1, N - 1 and N) and, for every pair, a count c in
max(0, a + b - N) .. min(a, b) -- uniform, at the independence point ab/N
or at either end.  The rows themselves are not needed: the r^2 is a
function of (a, b, c, N).  Only this container has the reference checkout;
the fixture is committed.  Usage: python tests/golden/make_wide_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

SIZES = [(1 << 24), (1 << 24) + 1, (1 << 24) + 12345, 22554 * 1000 + 7, (1 << 26),
         (1 << 30) + 3, (1 << 31) - 1]


def counts_for(rng, a, b, N, kind):
    lo = np.maximum(0, a + b - N)
    hi = np.minimum(a, b)
    if kind == 0:
        return lo + (rng.random(lo.shape) * (hi - lo + 1)).astype(np.int64)
    if kind == 1:
        c = np.rint(a.astype(np.float64) * b / N).astype(np.int64) + rng.integers(-2, 3, lo.shape)
        return np.clip(c, lo, hi)
    return np.where(rng.random(lo.shape) < 0.5, lo, hi)


def main():
    rng = np.random.default_rng(2024)
    out = {"sizes": np.array(SIZES, np.int64)}
    for k, N in enumerate(SIZES):
        bcA = rng.integers(0, N + 1, 24, dtype=np.int64)
        bcB = rng.integers(0, N + 1, 24, dtype=np.int64)
        bcA[:4] = [0, 1, N - 1, N]
        bcB[:4] = [N, N - 1, 1, 0]
        # counts below 2^32 (uint32 counts of the engine)
        bcA = np.minimum(bcA, (1 << 32) - 1)
        bcB = np.minimum(bcB, (1 << 32) - 1)
        a = np.broadcast_to(bcA[None, :], (24, 24))
        b = np.broadcast_to(bcB[:, None], (24, 24))
        c = counts_for(rng, a, b, N, k % 3)
        r2 = oracle.ref_r2(bcA.astype(np.uint32), bcB.astype(np.uint32),
                           c.astype(np.uint64), N)
        out[f"bcA_{k}"] = bcA.astype(np.uint32)
        out[f"bcB_{k}"] = bcB.astype(np.uint32)
        out[f"counts_{k}"] = c.astype(np.uint32)
        out[f"r2_{k}"] = r2.astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "r2_wide.npz"), **out)
    print("wrote r2_wide.npz for N =", SIZES)


if __name__ == "__main__":
    main()
