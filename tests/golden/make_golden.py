"""Write the golden fixtures in this directory from the reference itself.

Runs the reference's own C sources, compiled where they lie by
``make -C oracle ref`` (oracle/_ref/libqldref.so), on small seeded inputs and
stores inputs and outputs as .npz data:

* ``pack_kat.npz``   -- compressSnip_x64 (snip_processing.c:783-815) on
                        character strings -> packed words + bit counts
* ``fprint_kat.npz`` -- fprint_f (fast_print.c:35-103) float -> text
* ``case_*.npz``     -- tables, counts from the reference's blocked popcount
                        gemm (gemm_cpu.c:280-341), r^2 from
                        get_pairwise_ld_score (gemm_cpu.c:343-404) and the
                        report text of the reference's correlate()
                        (gemm_cpu.c:420-623 -> correlate_IO.c:378-429)

Only this container has the reference checkout; the fixtures are committed
so tests elsewhere need neither the reference nor this script.
Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def pack_rows(bits: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """0/1 matrix (rows x N) -> packed rows through the reference packer."""
    rows = []
    bcs = []
    for r in bits:
        w, bc = oracle.ref_compress_snip(bytes((r + 48).astype(np.uint8)))
        rows.append(w)
        bcs.append(bc)
    return np.array(rows, dtype=np.uint64), np.array(bcs, dtype=np.uint32)


def haplotypes(rng, rows, N, p_lo=0.02, p_hi=0.5, copy=0.85):
    """Small LD-structured 0/1 matrix (Markov copying along rows)."""
    out = np.zeros((rows, N), dtype=np.int64)
    for s in range(rows):
        p = rng.uniform(p_lo, p_hi)
        fresh = (rng.random(N) < p).astype(np.int64)
        if s == 0 or rng.random() < 0.1:
            out[s] = fresh
        else:
            keep = rng.random(N) < copy
            out[s] = np.where(keep, out[s - 1], fresh)
    return out


def special_rows(rng, N, base):
    """Edge rows: all zero, all one, a duplicate, a complement, one bit."""
    one = np.zeros(N, dtype=np.int64)
    one[rng.integers(N)] = 1
    return np.stack([np.zeros(N, dtype=np.int64), np.ones(N, dtype=np.int64),
                     base[0].copy(), 1 - base[1], one])


def make_case(name, rng, N, m, n, posWset2, r2limit, specials=True, tmp="."):
    bitsA = haplotypes(rng, m, N)
    if specials and m >= 8:
        bitsA[-5:] = special_rows(rng, N, bitsA)
    if posWset2:
        bitsB = haplotypes(rng, n, N)
        if specials and n >= 8:
            bitsB[:3] = np.stack([bitsA[1], 1 - bitsA[2], bitsA[-1]])
    else:
        bitsB = bitsA
        n = m
    A, bcA = pack_rows(bitsA)
    B, bcB = pack_rows(bitsB)
    posA = np.sort(rng.choice(10**6, m, replace=False)).astype(np.uint32) + 1
    posB = posA if not posWset2 else (
        np.sort(rng.choice(10**6, n, replace=False)).astype(np.uint32) + 1)
    idA = np.array([f"rs{p}" for p in posA])
    idB = np.array([f"rs{p}" for p in posB])
    counts = oracle.ref_counts(A, B)
    r2 = oracle.ref_r2(bcA, bcB, counts, N)
    text = oracle.ref_correlate_report(os.path.join(tmp, name + ".txt"), A, posA,
                                       list(idA), bcA, B, posB, list(idB), bcB,
                                       N, r2limit, posWset2)
    np.savez_compressed(os.path.join(HERE, f"case_{name}.npz"),
                        A=A, bcA=bcA, B=B, bcB=bcB, posA=posA, posB=posB,
                        idA=idA, idB=idB, snp_size=np.int64(N),
                        posWset2=np.int64(posWset2), r2limit=np.float32(r2limit),
                        counts=counts.astype(np.uint32), r2=r2,
                        report=np.frombuffer(text.encode(), dtype=np.uint8))
    print(f"case {name}: N={N} m={m} n={n} posWset2={posWset2} "
          f"lines={text.count(chr(10)) - 1}")


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref/libqldref.so missing: run `make -C oracle ref`")
    rng = np.random.default_rng(20261203)
    tmp = tempfile.mkdtemp()

    # Packing known answers, including the three fixed vectors.
    strings = [b"1" + b"0" * 63, b"01" * 32, b"1" * 65, b"1", b"0", b"10" * 48]
    for n in (1, 2, 31, 32, 33, 63, 64, 65, 96, 127, 128, 129, 200):
        strings.append(bytes(rng.choice([48, 49], n).astype(np.uint8)))
    lens = np.array([len(s) for s in strings], dtype=np.int64)
    flat = np.frombuffer(b"".join(strings), dtype=np.uint8)
    packed = [oracle.ref_compress_snip(s) for s in strings]
    W = max(len(p[0]) for p in packed)
    words = np.zeros((len(strings), W), dtype=np.uint64)
    for k, (w, _) in enumerate(packed):
        words[k, :len(w)] = w
    bcs = np.array([p[1] for p in packed], dtype=np.uint32)
    np.savez_compressed(os.path.join(HERE, "pack_kat.npz"), lens=lens,
                        chars=flat, words=words, bc=bcs)

    # Float printer known answers.
    vals = [0.0, 0.5, 0.25, 1.0, 0.1, 0.2, 0.3, 1e-9, 1e-10, 0.999999940395,
            1.00000012, 1.0000999, 0.333333343, 12.5, 99.999, 0.000123]
    vals += list(rng.random(300).astype(np.float32))
    vals += list((rng.integers(1, 16622, 200) / np.float32(16622)).astype(np.float32))
    vals = np.array(vals, dtype=np.float32)
    path = os.path.join(tmp, "f.txt")
    texts = [oracle.ref_fprint_f(float(v), path) for v in vals]
    np.savez_compressed(os.path.join(HERE, "fprint_kat.npz"), vals=vals,
                        texts=np.array(texts))

    make_case("tri_n63", rng, 63, 37, 0, 0, 0.0, tmp=tmp)
    make_case("tri_n64", rng, 64, 40, 0, 0, 0.1, tmp=tmp)
    make_case("rect_n65", rng, 65, 29, 41, 1, 0.0, tmp=tmp)
    make_case("tri_n130", rng, 130, 70, 0, 0, 0.2, tmp=tmp)
    make_case("rect_n96", rng, 96, 40, 33, 1, 0.05, tmp=tmp)
    make_case("tri_n1000", rng, 1000, 150, 0, 0, 0.2, tmp=tmp)
    make_case("rect_n1", rng, 1, 6, 5, 1, 0.0, specials=False, tmp=tmp)
    make_case("tri_n16622", rng, 16622, 24, 0, 0, 0.0, tmp=tmp)


if __name__ == "__main__":
    main()
