"""Pin the CPU oracle: golden fixtures from the reference, known answers,
library routines (numpy matmul) and invariants.  CPU only."""
import math
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, golden_cases, load_case, pack_bits, unpack_bits


# ---- packing -----------------------------------------------------------------

def test_pack_fixed_vectors():
    w, bc = oracle.pack_chars(b"1" + b"0" * 63)
    assert w.tolist() == [0x8000000000000000] and bc == 1
    w, bc = oracle.pack_chars(b"01" * 32)
    assert w.tolist() == [0x5555555555555555] and bc == 32
    w, bc = oracle.pack_chars(b"1" * 65)
    assert w.tolist() == [0xFFFFFFFFFFFFFFFF, 0x8000000000000000] and bc == 65


def test_pack_matches_golden():
    z = np.load(os.path.join(GOLDEN, "pack_kat.npz"))
    off = 0
    for k, n in enumerate(z["lens"]):
        chars = z["chars"][off:off + n].tobytes()
        off += n
        w, bc = oracle.pack_chars(chars)
        assert w.tolist() == z["words"][k, :len(w)].tolist()
        assert bc == z["bc"][k]
        # numpy packbits (a library routine) agrees as well
        pw, pbc = pack_bits(np.frombuffer(chars, np.uint8)[None, :] & 1)
        assert pw[0].tolist() == w.tolist() and pbc[0] == bc


# ---- counts ------------------------------------------------------------------

@pytest.mark.parametrize("N,m,n", [(1, 3, 4), (63, 17, 9), (64, 8, 8), (65, 11, 13),
                                   (96, 5, 30), (130, 20, 7), (1000, 33, 21)])
def test_counts_naive_equals_matmul(N, m, n):
    rng = np.random.default_rng(N * 7 + m)
    UA = (rng.random((m, N)) < rng.random((m, 1))).astype(np.int64)
    UB = (rng.random((n, N)) < rng.random((n, 1))).astype(np.int64)
    A, _ = pack_bits(UA)
    B, _ = pack_bits(UB)
    assert (unpack_bits(A, N) == UA).all()
    ref = UB @ UA.T                     # out[i, j] = <B_i, A_j>
    assert (oracle.counts_naive(A, B) == ref).all()


@pytest.mark.parametrize("m,n,W", [(1, 1, 1), (5, 3, 2), (73, 241, 5), (300, 7, 257),
                                   (130, 130, 513), (245, 250, 260)])
def test_counts_blocked_equals_naive(m, n, W):
    rng = np.random.default_rng(m * 1000 + n + W)
    A = rng.integers(0, 2**63, (m, W), dtype=np.uint64) * 2 + rng.integers(0, 2, (m, W), dtype=np.uint64)
    B = rng.integers(0, 2**63, (n, W), dtype=np.uint64)
    ref = oracle.counts_naive(A, B)
    assert (oracle.counts_blocked(A, B, tri=False) == ref).all()
    if m == n:
        got = oracle.counts_blocked(A, tri=True)
        full = oracle.counts_naive(A)
        iu = np.triu_indices(m, 0)       # j >= i cells
        assert (got[iu] == full[iu]).all()


def test_count_invariants():
    import qld_amd
    N = 777
    A, bc = qld_amd.synth_blocks(3, 90, N)
    c = oracle.counts_naive(A)
    assert (np.diag(c) == bc).all()                 # c_ii = bc_i
    assert (c == c.T).all()                         # symmetry
    a = bc[None, :].astype(np.int64)
    b = bc[:, None].astype(np.int64)
    assert (c >= np.maximum(0, a + b - N)).all() and (c <= np.minimum(a, b)).all()
    # sample permutation invariance (same permutation on every row)
    U = unpack_bits(A, N)
    perm = np.random.default_rng(0).permutation(N)
    P, _ = pack_bits(U[:, perm])
    assert (oracle.counts_naive(P) == c).all()


# ---- r^2 ---------------------------------------------------------------------

def _r2_of_bits(a_bits, b_bits, mode):
    A, bcA = pack_bits(np.array([a_bits]))
    B, bcB = pack_bits(np.array([b_bits]))
    c = oracle.counts_naive(A, B)
    N = len(a_bits)
    return {"ref": oracle.r2_reference, "exact": oracle.r2_exact,
            "fp64": oracle.r2_fp64}[mode](bcA, bcB, c, N)[0, 0]


@pytest.mark.parametrize("mode", ["ref", "exact", "fp64"])
def test_r2_known_answers(mode):
    assert _r2_of_bits([1, 1, 0, 0], [1, 0, 1, 0], mode) == 0.0
    assert abs(_r2_of_bits([1, 1, 0, 0], [1, 0, 0, 0], mode) - 1.0 / 3.0) < 1e-7
    x = [1, 0, 1, 1, 0, 0, 1, 0, 1]
    assert abs(_r2_of_bits(x, x, mode) - 1.0) < 1e-6
    assert abs(_r2_of_bits(x, [1 - v for v in x], mode) - 1.0) < 1e-6
    assert _r2_of_bits([0] * 9, x, mode) == 0.0          # bc == 0 -> 0
    assert math.isnan(_r2_of_bits([1] * 9, x, mode))     # bc == N -> 0/0


def test_r2_closed_form_brute_force():
    """fp64 closed form == r^2 from the 2x2 haplotype table by definition."""
    rng = np.random.default_rng(5)
    for _ in range(50):
        N = int(rng.integers(4, 200))
        a = (rng.random(N) < rng.random()).astype(int)
        b = (rng.random(N) < rng.random()).astype(int)
        if a.sum() in (0, N) or b.sum() in (0, N):
            continue
        pA, pB, pAB = a.mean(), b.mean(), (a & b).mean()
        D = pAB - pA * pB
        r2 = D * D / (pA * (1 - pA) * pB * (1 - pB))
        # also equals the squared Pearson correlation of the 0/1 vectors
        assert abs(r2 - np.corrcoef(a, b)[0, 1] ** 2) < 1e-12
        assert abs(_r2_of_bits(a, b, "fp64") - r2) < 1e-12


@pytest.mark.parametrize("N", [63, 16622, 100000])
def test_r2_modes_within_tolerance_of_fp64(N):
    import qld_amd
    A, bc = qld_amd.synth_blocks(11, 120, N)
    c = oracle.counts_naive(A)
    r64 = oracle.r2_fp64(bc, bc, c, N)
    ok = np.isfinite(r64)
    for f in (oracle.r2_reference, oracle.r2_exact):
        r = f(bc, bc, c, N).astype(np.float64)
        assert np.nanmax(np.abs(r - r64)[ok]) <= 1e-6
    rex = oracle.r2_exact(bc, bc, c, N).astype(np.float64)
    # exact mode: within one float ulp of the FP64 value
    ulp = np.spacing(np.abs(r64[ok]).astype(np.float32)).astype(np.float64)
    assert (np.abs(rex[ok] - r64[ok]) <= ulp).all()


# ---- golden fixtures ---------------------------------------------------------

@pytest.mark.parametrize("path", golden_cases(), ids=lambda p: os.path.basename(p))
def test_oracle_matches_golden(path, tmp_path):
    d = load_case(path)
    N = int(d["snp_size"])
    tri = int(d["posWset2"]) == 0
    cnt = oracle.counts_naive(d["A"], d["B"])
    assert (cnt == d["counts"]).all()
    r2 = oracle.r2_reference(d["bcA"], d["bcB"], cnt, N, tri=False)
    assert (r2.view(np.uint32) == d["r2"].view(np.uint32)).all()
    blk = oracle.counts_blocked(d["A"], None if tri else d["B"], tri=tri)
    sel = np.triu_indices(cnt.shape[0], 1) if tri else np.nonzero(np.ones_like(cnt))
    assert (blk[sel] == cnt[sel]).all()
    text = oracle.write_results_text(tmp_path / "r.txt", d["posA"], list(d["idA"]),
                                     d["bcA"], d["posB"], list(d["idB"]), d["bcB"],
                                     N, d["r2"], float(d["r2limit"]), int(d["posWset2"]))
    assert text == d["report"]


def test_golden_vectors_catch_fused_numerator():
    """A plausible mistake -- fusing p12 - p1*p2 into one FMA -- changes at
    least one r^2 bit pattern of the fixtures, so they pin that rounding."""
    diffs = 0
    for path in golden_cases():
        d = load_case(path)
        N = np.float32(int(d["snp_size"]))
        c = d["counts"].astype(np.float64)
        p1 = (d["bcA"].astype(np.float32) / N)[None, :].astype(np.float64)
        p2 = (d["bcB"].astype(np.float32) / N)[:, None].astype(np.float64)
        p12 = (c.astype(np.float32) / N).astype(np.float64)
        t = (p12 - p1 * p2).astype(np.float32)      # single rounding (FMA)
        x = (p1 * p2).astype(np.float32).astype(np.float64)
        den = (x * (1.0 - p1)) * (1.0 - p2)
        with np.errstate(all="ignore"):
            r = ((t * t).astype(np.float64) / den).astype(np.float32)
        ok = (d["bcA"][None, :] > 0) & (d["bcB"][:, None] > 0) & np.isfinite(r)
        diffs += int((r.view(np.uint32) != d["r2"].view(np.uint32))[ok].sum())
    assert diffs > 0


def test_fprint_matches_golden():
    z = np.load(os.path.join(GOLDEN, "fprint_kat.npz"))
    for v, t in zip(z["vals"], z["texts"]):
        assert oracle.fprint_f(float(v)) == str(t)
    assert oracle.fprint_f(0.5) == "0.500000000"
    assert oracle.fprint_f(0.25) == "0.250000000"
    assert oracle.fprint_f(1.0) == "1."


# ---- live comparison with the reference build (this container only) --------

ref_only = pytest.mark.skipif(not oracle.ref_available(),
                              reason="reference checkout not built here")


@ref_only
@pytest.mark.parametrize("N,m,n", [(16622, 40, 33), (100000, 9, 12), (127, 60, 60)])
def test_oracle_matches_reference_build(N, m, n):
    import qld_amd
    A, bcA, B, bcB = qld_amd.synth_regions(99, m, n, N)
    cr = oracle.ref_counts(A, B)
    co = oracle.counts_naive(A, B)
    assert (cr == co).all()
    rr = oracle.ref_r2(bcA, bcB, cr, N)
    ro = oracle.r2_reference(bcA, bcB, co, N)
    assert (rr.view(np.uint32) == ro.view(np.uint32)).all()


# ---- sample counts of 2^24 and more (the popcount engine's wide jobs) --------

def test_r2_wide_sample_counts_match_reference_fixture():
    """N from 2^24 to 2^31 - 1, where the reference's float casts of the
    counts and of N round: the oracle's r^2 equals get_pairwise_ld_score's
    bit for bit (tests/golden/r2_wide.npz, written by the reference build),
    and the exact form stays within 1e-6 of the FP64 closed form."""
    with np.load(os.path.join(GOLDEN, "r2_wide.npz"), allow_pickle=False) as z:
        sizes = z["sizes"]
        for k, N in enumerate(sizes):
            bcA, bcB, c, want = z[f"bcA_{k}"], z[f"bcB_{k}"], z[f"counts_{k}"], z[f"r2_{k}"]
            got = oracle.r2_reference(bcA, bcB, c, int(N))
            assert (got.view(np.uint32) == want.view(np.uint32)).all(), f"N={N}"
            ex = oracle.r2_exact(bcA, bcB, c, int(N))
            r64 = oracle.r2_fp64(bcA, bcB, c, int(N))
            ok = np.isfinite(r64)
            assert np.abs(ex[ok].astype(np.float64) - r64[ok]).max() <= 1e-6, f"N={N}"
