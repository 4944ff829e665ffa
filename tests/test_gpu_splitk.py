"""Split-K plain calls (128-tiles of the direct fp4 engine): each tile's word
groups summed by several workgroups, the last to finish adding the others'
partial sums and writing the tile.  Checked against the oracle through the
C ABI, to the same bar as the unsplit path: counts bit-exact, REFERENCE r^2
bit-identical to the reference's formula, EXACT r^2 bit-identical to the
oracle's exact restatement.  Needs an MI355X."""
import sys

import numpy as np
import pytest

import oracle
import qld_amd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check(res, cnt, A, bcA, B, bcB, N, tri, mode):
    n, m = res.shape
    sel = np.triu(np.ones((n, m), dtype=bool), 1) if tri else np.ones((n, m), dtype=bool)
    want_c = oracle.counts_naive(A, B)
    assert (cnt[sel] == want_c[sel]).all(), "counts differ"
    f = oracle.r2_reference if mode == "reference" else oracle.r2_exact
    want = f(bcA, bcB, want_c, N)
    bad = np.nonzero(res.view(np.uint32)[sel] != want.view(np.uint32)[sel])[0]
    assert bad.size == 0, f"{bad.size} r2 cells differ"
    if tri:
        assert (res[~sel] == 0).all() and (cnt[~sel] == 0).all()


# (kind, N, shape): 65, 4, 391 and 782 word groups; tiles on the diagonal,
# ragged edges, a single row; chunk counts above the group count leave empty
# chunks; at 782 groups 2 and 3 chunks keep fp32 partial sums (more than 255
# groups per chunk), more chunks 16-bit ones
CASES = [
    ("tri", 16622, 300), ("tri", 16622, 129), ("tri", 16622, 1),
    ("rect", 16622, (129, 127)), ("rect", 1000, (300, 257)), ("rect", 100000, (65, 200)),
    ("tri", 100000, 257), ("rect", 200000, (40, 150)),
]


@pytest.mark.parametrize("mode", ["reference", "exact"])
@pytest.mark.parametrize("chunks", [2, 3, 5, 8, 64])
def test_forced_split(mode, chunks, engine_factory):
    e = engine_factory(engine="mfma_fp4", r2_mode=mode)
    e.set_ksplit(chunks)
    for kind, N, shape in CASES:
        say("split", mode, chunks, kind, N, shape)
        if kind == "tri":
            A, bc = qld_amd.synth_blocks(500 + shape + N, shape, N)
            res, cnt = e.correlate(A, bc, snp_size=N, triangular=True, counts=True)
            _check(res, cnt, A, bc, A, bc, N, True, mode)
        else:
            m, n = shape
            A, bA, B, bB = qld_amd.synth_regions(11 + m * n + N, m, n, N)
            res, cnt = e.correlate(A, bA, B, bB, snp_size=N, triangular=False, counts=True)
            _check(res, cnt, A, bA, B, bB, N, False, mode)
        assert e.last_ksplit == chunks


def test_split_off_and_auto(engine_factory):
    """ksplit 1 never splits; by shape, the 2,000-SNP triangle at 16,622
    samples (136 tiles) is split, a 12,000-SNP one (256-tiles) is not, and a
    job of 4-group rows is not (chunks keep at least 4 groups)."""
    e = engine_factory(engine="mfma_fp4", r2_mode="reference")
    A, bc = qld_amd.synth_blocks(20261204, 2000, 16622)
    say("auto 2000")
    res, cnt = e.correlate(A, bc, snp_size=16622, triangular=True, counts=True)
    assert e.last_ksplit >= 2
    _check(res, cnt, A, bc, A, bc, 16622, True, "reference")
    e.set_ksplit(1)
    res1, cnt1 = e.correlate(A, bc, snp_size=16622, triangular=True, counts=True)
    assert e.last_ksplit == 1
    assert (res1.view(np.uint32) == res.view(np.uint32)).all() and (cnt1 == cnt).all()
    e.set_ksplit(0)
    A2, bc2 = qld_amd.synth_blocks(5, 300, 1000)
    e.correlate(A2, bc2, snp_size=1000, triangular=True)
    assert e.last_ksplit == 1
    A3, bc3 = qld_amd.synth_blocks(6, 12000, 64 * 20)
    e.correlate(A3, bc3, snp_size=64 * 20, triangular=True)
    assert e.last_ksplit == 1
    for bad in (-1, 65):
        with pytest.raises(qld_amd.QldError):
            e.set_ksplit(bad)


def test_counters_reset_between_calls(engine_factory):
    """The arrival counters are left zero by every launch: calls of growing,
    shrinking and repeated shapes and chunk counts on one context agree with
    the oracle and with themselves."""
    e = engine_factory(engine="mfma_fp4", r2_mode="exact")
    N = 16622
    A, bc = qld_amd.synth_blocks(9, 700, N)
    first = None
    for chunks, m in [(8, 300), (3, 700), (8, 300), (5, 129), (3, 700), (8, 300)]:
        e.set_ksplit(chunks)
        say("sequence", chunks, m)
        res, cnt = e.correlate(A[:m], bc[:m], snp_size=N, triangular=True, counts=True)
        assert e.last_ksplit == chunks
        _check(res, cnt, A[:m], bc[:m], A[:m], bc[:m], N, True, "exact")
        if m == 300:
            if first is None:
                first = res.copy()
            assert (res.view(np.uint32) == first.view(np.uint32)).all()


def test_results_or_counts_only(engine_factory):
    e = engine_factory(engine="mfma_fp4", r2_mode="reference")
    e.set_ksplit(4)
    N = 16622
    A, bA, B, bB = qld_amd.synth_regions(12, 200, 150, N)
    want_c = oracle.counts_naive(A, B)
    say("counts only")
    _, cnt = e.correlate(A, bA, B, bB, snp_size=N, triangular=False, results=False, counts=True)
    assert (cnt == want_c).all()
    say("results only")
    res, _ = e.correlate(A, bA, B, bB, snp_size=N, triangular=False)
    want = oracle.r2_reference(bA, bB, want_c, N)
    assert (res.view(np.uint32) == want.view(np.uint32)).all()


def test_device_rows_sampled(engine_factory):
    """The 2,000-SNP x 16,622-sample triangle through the device-buffer entry
    point (split by shape), every row checked against the oracle."""
    N, m = 16622, 2000
    A, bc = qld_amd.synth_blocks(20261203 + 1, m, N)
    W = A.shape[1]
    e = engine_factory(engine="mfma_fp4", r2_mode="reference")
    dA = e.alloc(A.nbytes).upload(A)
    dbc = e.alloc(bc.nbytes).upload(bc)
    dres = e.alloc(m * m * 4)
    say("device 2000")
    e.correlate_dev(dA, dbc, m, dA, dbc, m, W, N, True, dres, m, None, 0)
    e.sync()
    assert e.last_ksplit >= 2
    res = dres.download((m, m), np.float32)
    want = oracle.r2_reference(bc, bc, oracle.counts_naive(A), N)
    sel = np.triu(np.ones((m, m), dtype=bool), 1)
    assert (res.view(np.uint32)[sel] == want.view(np.uint32)[sel]).all()
    for b in (dA, dbc, dres):
        b.free()


@pytest.mark.parametrize("groups, chunks", [(510, 2), (512, 2), (765, 3), (768, 3)])
def test_partial_sum_width_bounds(groups, chunks, engine_factory):
    """Partial sums are 16-bit while a chunk holds at most 255 word groups
    (counts up to 65,280) and fp32 beyond: rows of all ones and their
    complements at chunk sizes 255 and 256 groups, counts against the
    oracle (a pair of all-one rows counts every sample)."""
    N = groups * 256
    W = groups * 4
    rng = np.random.default_rng(groups)
    bits = np.zeros((70, W * 64), dtype=np.uint8)
    bits[:10] = 1                                       # all ones
    bits[10:40] = rng.random((30, W * 64)) < 0.97       # dense
    bits[40:] = 1 - bits[10:40]                         # complements
    A = np.packbits(bits, axis=1, bitorder="big").view(">u8").astype(np.uint64)
    bc = bits.sum(axis=1).astype(np.uint32)
    e = engine_factory(engine="mfma_fp4", r2_mode="exact")
    e.set_ksplit(chunks)
    say("width", groups, chunks)
    res, cnt = e.correlate(A, bc, snp_size=N, triangular=True, counts=True)
    assert e.last_ksplit == chunks
    assert cnt[0, 1] == N
    _check(res, cnt, A, bc, A, bc, N, True, "exact")


def test_timing_switch(engine_factory):
    """qld_set_timing: with the per-call events off a call records no kernel
    time (asking for it is an error); on again, the next call does."""
    e = engine_factory(engine="mfma_fp4", r2_mode="exact", timing=True)
    A, bc = qld_amd.synth_blocks(3, 500, 16622)
    e.correlate(A, bc, snp_size=16622, triangular=True)
    assert e.last_contract_ms() > 0
    e.set_timing(False)
    e.correlate(A, bc, snp_size=16622, triangular=True)
    with pytest.raises(qld_amd.QldError):
        e.last_contract_ms()
    e.set_timing(True)
    e.correlate(A, bc, snp_size=16622, triangular=True)
    assert e.last_contract_ms() > 0
