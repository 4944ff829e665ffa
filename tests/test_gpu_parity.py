"""Parity of the HIP engine with the oracle and the reference's golden
fixtures, through the C ABI.  Needs an MI355X.

Bar: counts bit-exact; REFERENCE-mode r^2 bit-identical to the reference's
get_pairwise_ld_score; EXACT-mode r^2 bit-identical to the oracle's exact
restatement; both within 1e-6 absolute of the FP64 closed form."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import oracle
import qld_amd
from conftest import golden_cases, load_case

pytestmark = pytest.mark.gpu

ENGINES = ["popcount", "mfma_fp4", "mfma_fp4p", "mfma_fp4x", "mfma_i8"]
# Shapes large enough for 256 x 256 tiles on every CU (the AUTO engine path).
BIG_SHAPES = [(1100, 900), (600, 3000)]
TILES = [256, 128]
TOL_FP64 = 1e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _sel(n, m, tri):
    if tri:
        return np.triu(np.ones((n, m), dtype=bool), 1)
    return np.ones((n, m), dtype=bool)


def _check(res, cnt, A, bcA, B, bcB, N, tri, mode, ref_counts=None):
    n, m = res.shape
    sel = _sel(n, m, tri)
    want_c = ref_counts if ref_counts is not None else oracle.counts_naive(A, B)
    assert (cnt[sel] == want_c[sel]).all(), "counts differ"
    f = oracle.r2_reference if mode == "reference" else oracle.r2_exact
    want = f(bcA, bcB, want_c, N)
    got_bits, want_bits = res.view(np.uint32)[sel], want.view(np.uint32)[sel]
    bad = np.nonzero(got_bits != want_bits)[0]
    assert bad.size == 0, f"{bad.size} r2 cells differ, first {res[sel][bad[:3]]} vs {want[sel][bad[:3]]}"
    r64 = oracle.r2_fp64(bcA, bcB, want_c, N)[sel]
    ok = np.isfinite(r64)
    if ok.any():
        assert np.abs(res[sel][ok].astype(np.float64) - r64[ok]).max() <= TOL_FP64


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("path", golden_cases(), ids=lambda p: os.path.basename(p))
def test_golden_cases(engine, path, engine_factory):
    d = load_case(path)
    say("golden", d["name"], engine)
    N = int(d["snp_size"])
    tri = int(d["posWset2"]) == 0
    e = engine_factory(engine=engine, r2_mode="reference")
    if tri:
        res, cnt = e.correlate(d["A"], d["bcA"], snp_size=N, triangular=True, counts=True)
    else:
        res, cnt = e.correlate(d["A"], d["bcA"], d["B"], d["bcB"], snp_size=N,
                               triangular=False, counts=True)
    sel = _sel(*res.shape, tri)
    assert (cnt[sel] == d["counts"][sel]).all()
    assert (res.view(np.uint32)[sel] == d["r2"].view(np.uint32)[sel]).all()
    if tri:   # cells outside the triangle come back zeroed
        assert (res[~sel] == 0).all() and (cnt[~sel] == 0).all()


@pytest.mark.parametrize("path", golden_cases(), ids=lambda p: os.path.basename(p))
def test_dropin_report_text(path, tmp_path, engine_factory):
    """qld_correlate_report (the correlate_gpu drop-in) writes the reference's
    report byte for byte."""
    d = load_case(path)
    say("report", d["name"])
    e = engine_factory(engine="auto", r2_mode="reference")
    lib = qld_amd.host_lib()
    libc = oracle.libc()
    m, n = len(d["posA"]), len(d["posB"])
    tri = int(d["posWset2"]) == 0
    ida = (C.c_char_p * m)(*[s.encode() for s in d["idA"]])
    idb = (C.c_char_p * n)(*[s.encode() for s in d["idB"]])
    A = np.ascontiguousarray(d["A"])
    B = A if tri else np.ascontiguousarray(d["B"])
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    posB = d["posA"] if tri else d["posB"]
    bcB = d["bcA"] if tri else d["bcB"]
    out = tmp_path / "r.txt"
    fp = libc.fopen(str(out).encode(), b"w")
    rc = lib.qld_correlate_report(e._ctx, fp, p(A), p(d["posA"]), C.cast(ida, C.c_void_p),
                                  p(d["bcA"]), m, p(B), p(np.ascontiguousarray(posB)),
                                  C.cast(idb if not tri else ida, C.c_void_p),
                                  p(np.ascontiguousarray(bcB)), n, A.shape[1],
                                  int(d["snp_size"]), int(d["posWset2"]),
                                  C.c_float(float(d["r2limit"])))
    libc.fclose(fp)
    assert rc == 0, qld_amd.hip_lib().qld_gpu_last_error(e._ctx)
    assert out.read_text() == d["report"]


SHAPES_TRI = [1, 2, 127, 128, 129, 300]
SHAPES_RECT = [(1, 1), (5, 300), (129, 127), (300, 257)]
SAMPLES = [1, 63, 64, 65, 96, 130, 16622, 100000]


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("mode", ["reference", "exact"])
@pytest.mark.parametrize("N", SAMPLES)
def test_shapes_triangular(engine, mode, N, engine_factory):
    e = engine_factory(engine=engine, r2_mode=mode)
    for m in SHAPES_TRI:
        say("tri", engine, mode, N, m)
        A, bc = qld_amd.synth_blocks(1000 + m + N, m, N)
        res, cnt = e.correlate(A, bc, snp_size=N, triangular=True, counts=True)
        _check(res, cnt, A, bc, A, bc, N, True, mode)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("mode", ["reference", "exact"])
@pytest.mark.parametrize("N", SAMPLES)
def test_shapes_rectangular(engine, mode, N, engine_factory):
    e = engine_factory(engine=engine, r2_mode=mode)
    for m, n in SHAPES_RECT:
        say("rect", engine, mode, N, m, n)
        A, bA, B, bB = qld_amd.synth_regions(7 + m * n + N, m, n, N)
        res, cnt = e.correlate(A, bA, B, bB, snp_size=N, triangular=False, counts=True)
        _check(res, cnt, A, bA, B, bB, N, False, mode)


@pytest.mark.parametrize("engine", ENGINES)
def test_monomorphic_and_duplicate_rows(engine, engine_factory):
    N = 16622
    A, bc = qld_amd.synth_blocks(77, 200, N)
    A[10] = 0
    A[11] = A[12]
    W = A.shape[1]
    A[13] = np.uint64(0xFFFFFFFFFFFFFFFF)
    A[13, W - 1] = np.uint64(((1 << (N % 64)) - 1) << (64 - N % 64))
    A[14] = A[15] ^ A[13]                                  # complement
    bc = np.array([int(sum(bin(int(w)).count("1") for w in r)) for r in A], np.uint32)
    for mode in ("reference", "exact"):
        e = engine_factory(engine=engine, r2_mode=mode)
        res, cnt = e.correlate(A, bc, snp_size=N, triangular=True, counts=True)
        _check(res, cnt, A, bc, A, bc, N, True, mode)
        assert res[10, 11] == 0.0 and np.isnan(res[13, 20])
        assert res[11, 12] == pytest.approx(1.0, abs=1e-6)
        assert res[14, 15] == pytest.approx(1.0, abs=1e-6)


@pytest.mark.parametrize("engine", ENGINES)
def test_full_size_sampled_rows(engine, engine_factory):
    """30,000 SNPs x 16,622 samples (the single-region benchmark shape):
    device-resident run, then 64 random rows checked against the oracle."""
    N, m = 16622, 30000
    A, bc = qld_amd.synth_blocks(20261203 + 2, m, N)
    W = A.shape[1]
    e = engine_factory(engine=engine, r2_mode="reference")
    dA = e.alloc(A.nbytes).upload(A)
    dbc = e.alloc(bc.nbytes).upload(bc)
    dres = e.alloc(m * m * 4)
    dcnt = e.alloc(m * m * 4)
    say("full size", engine)
    e.correlate_dev(dA, dbc, m, dA, dbc, m, W, N, True, dres, m, dcnt, m)
    e.sync()
    rows = np.sort(np.random.default_rng(3).choice(m, 64, replace=False))
    lib = qld_amd.hip_lib()
    for i in rows:
        r = np.empty(m, np.float32)
        c = np.empty(m, np.uint32)
        off = int(i) * m * 4
        assert lib.qld_d2h(e._ctx, r.ctypes.data_as(C.c_void_p), dres.ptr + off, m * 4) == 0
        assert lib.qld_d2h(e._ctx, c.ctypes.data_as(C.c_void_p), dcnt.ptr + off, m * 4) == 0
        want_c = oracle.counts_naive(A, A[i:i + 1])[0]
        j = np.arange(i + 1, m)
        assert (c[j] == want_c[j]).all()
        want = oracle.r2_reference(bc, bc[i:i + 1], want_c[None, :], N)[0]
        assert (r[j].view(np.uint32) == want[j].view(np.uint32)).all()
    for b in (dA, dbc, dres, dcnt):
        b.free()


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("mode", ["reference", "exact"])
def test_largest_sample_count(engine, mode, engine_factory):
    """N = 2^24 - 1, the largest sample count the MFMA engines accept (counts
    near 2^24 must stay exact in fp32 sums): dense, complementary, sparse and
    all-one rows, a triangle and a rectangle; N = 2^24 is rejected by the MFMA
    engines and taken by the popcount engine."""
    N = (1 << 24) - 1
    W = (N + 63) // 64
    rng = np.random.default_rng(24)
    bits = np.zeros((6, W * 64), dtype=np.uint8)
    bits[0, :N] = 1                                     # all ones (bc = N -> NaN r2)
    bits[1, :N] = rng.random(N) < 0.9                   # dense
    bits[2, :N] = 1 - bits[1, :N]                       # its complement
    bits[3, :N] = bits[1, :N]                           # a copy
    bits[3, rng.integers(0, N, 1000)] ^= 1
    bits[4, :N] = rng.random(N) < 0.001                 # sparse
    bits[5, :N] = rng.random(N) < 0.5
    A = np.packbits(bits, axis=1, bitorder="big").view(">u8").astype(np.uint64)
    bc = bits.sum(axis=1).astype(np.uint32)
    e = engine_factory(engine=engine, r2_mode=mode)
    say("max N", engine, mode)
    res, cnt = e.correlate(A, bc, snp_size=N, triangular=True, counts=True)
    _check(res, cnt, A, bc, A, bc, N, True, mode)
    res, cnt = e.correlate(A[:2], bc[:2], A[2:], bc[2:], snp_size=N, triangular=False,
                           counts=True)
    _check(res, cnt, A[:2], bc[:2], A[2:], bc[2:], N, False, mode)
    if engine == "popcount":        # takes any int sample count
        e.correlate(np.zeros((2, W + 1), np.uint64), np.zeros(2, np.uint32), snp_size=1 << 24)
    else:
        with pytest.raises(qld_amd.QldError) as x:
            e.correlate(np.zeros((2, W + 1), np.uint64), np.zeros(2, np.uint32),
                        snp_size=1 << 24)
        assert x.value.code == qld_amd.QLD_ERR_INVALID


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("mode", ["reference", "exact"])
def test_sample_count_above_row_bits(engine, mode, engine_factory):
    """With a sample list the reference packs only the listed samples but
    still divides by the header's sample count (snp_size = 22,554 for rows
    of 5,638 bits, as in its own test data): triangle, rectangle and the
    selection path against the oracle."""
    N, bits = 22554, 5638
    A, bc = qld_amd.synth_blocks(31, 300, bits)
    B, bB = qld_amd.synth_blocks(32, 170, bits)
    e = engine_factory(engine=engine, r2_mode=mode)
    say("N above row bits", engine, mode)
    res, cnt = e.correlate(A, bc, snp_size=N, triangular=True, counts=True)
    _check(res, cnt, A, bc, A, bc, N, True, mode)
    res, cnt = e.correlate(A, bc, B, bB, snp_size=N, triangular=False, counts=True)
    _check(res, cnt, A, bc, B, bB, N, False, mode)
    offs, cols, r2 = e.correlate_hits(A, bc, snp_size=N, upper=True, r2limit=0.05)
    full, _ = e.correlate(A, bc, snp_size=N, triangular=True)
    for i in range(A.shape[0]):
        c = cols[offs[i]:offs[i + 1]]
        j = np.arange(i + 1, A.shape[0])
        want = j[(full[i, j] >= 0.05) & (full[i, j] < 100)]
        assert (c == want).all()
        assert (r2[offs[i]:offs[i + 1]].view(np.uint32) == full[i, want].view(np.uint32)).all()


@pytest.mark.parametrize("engine", ENGINES)
def test_empty_inputs(engine, engine_factory):
    """Zero rows on either side: no work, no error, empty outputs; the
    context stays usable."""
    N = 16622
    A, bc = qld_amd.synth_blocks(11, 40, N)
    E = np.zeros((0, A.shape[1]), np.uint64)
    bE = np.zeros(0, np.uint32)
    e = engine_factory(engine=engine, r2_mode="reference")
    res, cnt = e.correlate(E, bE, snp_size=N, triangular=True, counts=True)
    assert res.shape == (0, 0) and cnt.shape == (0, 0)
    res, _ = e.correlate(A, bc, E, bE, snp_size=N, triangular=False)
    assert res.shape == (0, 40)
    res, _ = e.correlate(E, bE, A, bc, snp_size=N, triangular=False)
    assert res.shape == (40, 0)
    offs, cols, r2 = e.correlate_hits(A, bc, E, bE, snp_size=N, upper=False)
    assert len(cols) == 0 and len(r2) == 0 and int(offs[-1]) == 0
    res, cnt = e.correlate(A, bc, snp_size=N, triangular=True, counts=True)
    _check(res, cnt, A, bc, A, bc, N, True, "reference")


def test_contexts_on_concurrent_threads(engine_factory):
    """One context per calling thread, four threads at once (ctypes releases
    the GIL during each call): every thread's results stay bit-exact."""
    import threading
    N = 16622
    jobs = []
    for k, (m, eng) in enumerate([(1200, "auto"), (700, "mfma_fp4"), (300, "popcount"),
                                  (1100, "mfma_fp4x")]):
        A, bc = qld_amd.synth_blocks(900 + k, m, N)
        want_c = oracle.counts_naive(A)
        jobs.append((A, bc, want_c, oracle.r2_reference(bc, bc, want_c, N),
                     engine_factory(engine=eng, r2_mode="reference")))
    errors = []

    def work(A, bc, want_c, want, e):
        try:
            sel = np.triu(np.ones((A.shape[0],) * 2, dtype=bool), 1)
            for _ in range(3):
                res, cnt = e.correlate(A, bc, snp_size=N, triangular=True, counts=True)
                assert (cnt[sel] == want_c[sel]).all(), "counts differ"
                assert (res.view(np.uint32)[sel] == want.view(np.uint32)[sel]).all(), "r2 differs"
        except Exception as x:          # reported on the main thread
            errors.append(repr(x))

    threads = [threading.Thread(target=work, args=j) for j in jobs]
    say("threads", len(threads))
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in threads), "a thread did not finish"
    assert not errors, errors


def test_errors_are_status_codes(engine_factory):
    e = engine_factory(engine="popcount")
    A, bc = qld_amd.synth_blocks(1, 10, 100)
    B, bB = qld_amd.synth_blocks(2, 11, 100)
    with pytest.raises(qld_amd.QldError) as x:
        e.correlate(A, bc, B, bB, snp_size=100, triangular=True)
    assert x.value.code == qld_amd.QLD_ERR_INVALID
    for bad_n in (0, -5):
        with pytest.raises(qld_amd.QldError) as x:
            e.correlate(A, bc, snp_size=bad_n)
        assert x.value.code == qld_amd.QLD_ERR_INVALID
    with pytest.raises(qld_amd.QldError) as x:     # beyond the MFMA engines' exact range
        engine_factory(engine="mfma_fp4").correlate(A, bc, snp_size=1 << 24)
    assert x.value.code == qld_amd.QLD_ERR_INVALID
    small = engine_factory(engine="popcount", hbm_budget=1 << 10)
    with pytest.raises(qld_amd.QldError) as x:
        small.alloc(1 << 20)
    assert x.value.code == qld_amd.QLD_ERR_BUDGET
    # the context stays usable after errors
    res, _ = e.correlate(A, bc, snp_size=100)
    assert res.shape == (10, 10)


@pytest.mark.parametrize("engine", ["mfma_fp4", "mfma_fp4x", "mfma_i8"])
@pytest.mark.parametrize("mode", ["reference", "exact"])
@pytest.mark.parametrize("tile", TILES)
def test_forced_tile_shapes(engine, mode, tile, engine_factory):
    """Both tile edges on ragged rectangles and a triangle."""
    e = engine_factory(engine=engine, r2_mode=mode)
    e.set_plain_tile(tile)
    for N in (16622, 100000):
        for m, n in BIG_SHAPES:
            say("rect", engine, mode, tile, N, m, n)
            A, bA, B, bB = qld_amd.synth_regions(31 + m + N, m, n, N)
            res, cnt = e.correlate(A, bA, B, bB, snp_size=N, triangular=False, counts=True)
            _check(res, cnt, A, bA, B, bB, N, False, mode)
    A, bc = qld_amd.synth_blocks(8, 4200, 16622)
    res, cnt = e.correlate(A, bc, snp_size=16622, triangular=True, counts=True)
    rows = np.random.default_rng(0).choice(4200, 40, replace=False)
    want_c = oracle.counts_naive(A, A[rows])
    for k, i in enumerate(rows):
        j = np.arange(i + 1, 4200)
        assert (cnt[i, j] == want_c[k, j]).all()


@pytest.mark.parametrize("engine", ["mfma_i8", "mfma_fp4x", "mfma_fp4", "popcount"])
def test_widest_rows_many_rows(engine, engine_factory):
    """Rows of the largest accepted width, 262,144 words (N = 2^24 - 1), 300
    of them: counts up to 2^24 in the fp32 sums, and expanded operand rows
    of 8 MiB (fp4) / 16 MiB (int8), whose 256-row tiles would span 2^32
    bytes -- such launches drop to 128-tiles.  Triangle with either tile
    edge requested, a rectangle, and a window plan with 256-tiles asked for;
    every count against the oracle.  A width of 262,145 words is rejected by
    the MFMA engines (the popcount engine takes it)."""
    N = (1 << 24) - 1
    W = (N + 63) // 64
    rows = 300
    rng = np.random.default_rng(262144)
    A = rng.integers(0, 2**63, (rows, W), dtype=np.uint64) | (rng.integers(0, 2, (rows, W),
                                                                          dtype=np.uint64) << 63)
    A[:, -1] &= ~np.uint64(1)            # the pad bit of the last word (N = 64 W - 1)
    A[0, :] = ~np.uint64(0)              # an all-ones row: counts of 2^24 - 1
    A[0, -1] = ~np.uint64(1)
    bc = np.bitwise_count(A).sum(axis=1).astype(np.uint32)
    want = oracle.counts_blocked(A, tri=False)
    e = engine_factory(engine=engine, r2_mode="exact")
    for tile in (256, 128):
        if engine != "popcount":
            e.set_plain_tile(tile)
        say("widest rows", engine, tile)
        res, cnt = e.correlate(A, bc, snp_size=N, triangular=True, counts=True)
        sel = np.triu(np.ones((rows, rows), dtype=bool), 1)
        assert (cnt[sel] == want[sel]).all(), f"tile {tile}: counts differ"
    res, cnt = e.correlate(A[:130], bc[:130], A[130:], bc[130:], snp_size=N, triangular=False,
                           counts=True)
    assert (cnt == want[130:, :130]).all(), "rectangle counts differ"
    if engine != "popcount":
        e.set_tile_cutoff(0)
        t = e.table(A, bc, N)
        wins = qld_amd.window_layout(np.array([0, 20], np.int64), np.array([300, 270], np.int64))
        total = int((wins["size"].astype(np.int64) ** 2).sum())
        out = e.alloc(total * 4)
        plan = e.plan_windows(t, wins, out, qld_amd.BATCH_PERSISTENT)
        plan.run()
        e.sync()
        got = out.download(total, np.float32)
        for w in wins:
            s, n, off = int(w["start"]), int(w["size"]), int(w["out_offset"])
            g = got[off:off + n * n].reshape(n, n)
            want_r2 = oracle.r2_exact(bc[s:s + n], bc[s:s + n], want[s:s + n, s:s + n], N)
            sel = np.triu(np.ones((n, n), dtype=bool), 1)
            assert (g.view(np.uint32)[sel] == want_r2.view(np.uint32)[sel]).all(), \
                f"window {s}+{n} differs"
        plan.free()
        t.free()
        out.free()
    if engine == "popcount":        # uint32 counts: wider rows are taken
        e.correlate(np.zeros((2, W + 1), np.uint64), np.zeros(2, np.uint32), snp_size=1000)
    else:
        with pytest.raises(qld_amd.QldError) as x:
            e.correlate(np.zeros((2, W + 1), np.uint64), np.zeros(2, np.uint32), snp_size=1000)
        assert x.value.code == qld_amd.QLD_ERR_INVALID


@pytest.mark.parametrize("engine", ["popcount", "auto"])
@pytest.mark.parametrize("mode", ["reference", "exact"])
def test_wide_sample_counts(engine, mode, engine_factory):
    """Sample counts from 2^24 up, as the reference takes any int snp_size:
    they run on the popcount engine (AUTO picks it, uint32 counts, the
    reference's float casts of c and N in REFERENCE mode, 64-bit integer
    numerators in EXACT mode).  Rows of exactly N bits at N = 2^24 and
    2^24 + 12,345, rows of 2^26 bits (2^20 words, wider than the MFMA
    engines' 2^18), and N = 2^30 + 3 over rows of 5,000 bits (a sample list
    far below the header's count); all-one, dense, complementary, sparse
    rows; a triangle and a rectangle, counts and r^2 bit-exact against the
    oracle (pinned for these N by tests/golden/r2_wide.npz)."""
    rng = np.random.default_rng(26)
    for N, nbits in [(1 << 24, 1 << 24), ((1 << 24) + 12345, (1 << 24) + 12345),
                     (1 << 26, 1 << 26), ((1 << 30) + 3, 5000)]:
        W = (nbits + 63) // 64
        bits = np.zeros((6, W * 64), dtype=np.uint8)
        bits[0, :nbits] = 1
        bits[1, :nbits] = rng.random(nbits) < 0.9
        bits[2, :nbits] = 1 - bits[1, :nbits]
        bits[3, :nbits] = bits[1, :nbits]
        bits[3, rng.integers(0, nbits, 1000)] ^= 1
        bits[4, :nbits] = rng.random(nbits) < 0.001
        bits[5, :nbits] = rng.random(nbits) < 0.5
        A = np.packbits(bits, axis=1, bitorder="big").view(">u8").astype(np.uint64)
        bc = bits.sum(axis=1).astype(np.uint32)
        del bits
        e = engine_factory(engine=engine, r2_mode=mode)
        say("wide N", engine, mode, N, W)
        res, cnt = e.correlate(A, bc, snp_size=N, triangular=True, counts=True)
        assert e.last_engine == "popcount"
        _check(res, cnt, A, bc, A, bc, N, True, mode)
        res, cnt = e.correlate(A[:2], bc[:2], A[2:], bc[2:], snp_size=N, triangular=False,
                               counts=True)
        _check(res, cnt, A[:2], bc[:2], A[2:], bc[2:], N, False, mode)
