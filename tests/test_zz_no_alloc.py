"""The allocation-free entry point (qld_correlate_dev_noalloc) against
qld_correlate_dev, through the C ABI.  Needs an MI355X.

Both calls write into buffers pre-filled with the same sentinel, so cells one
call writes and the other does not show up as differences; r^2 bits and
counts must be identical.  Each case makes two no-allocation calls in a row,
the second on a smaller input (the scratch keeps the first call's contents)."""
import sys

import numpy as np
import pytest

import qld_amd

pytestmark = pytest.mark.gpu

SENTINEL = 0x7FC00BAD     # a NaN that no kernel writes


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(eng, fn, dA, dbA, m, dB, dbB, n, W, N, tri, counts):
    res = eng.alloc(n * m * 4).fill(SENTINEL)
    cnt = eng.alloc(n * m * 4).fill(SENTINEL) if counts else None
    fn(dA, dbA, m, dB, dbB, n, W, N, tri, res, m, cnt, m)
    eng.sync()
    r = res.download((n, m), np.uint32)
    c = cnt.download((n, m), np.uint32) if counts else None
    res.free()
    if cnt:
        cnt.free()
    return r, c


def _case(eng, A, bcA, B, bcB, N, tri, counts):
    m, W = A.shape
    n = B.shape[0]
    dA, dbA = eng.alloc(A.nbytes).upload(A), eng.alloc(bcA.nbytes).upload(bcA)
    if tri:
        dB, dbB = dA, dbA
    else:
        dB, dbB = eng.alloc(B.nbytes).upload(B), eng.alloc(bcB.nbytes).upload(bcB)
    got = _run(eng, eng.correlate_dev_noalloc, dA, dbA, m, dB, dbB, n, W, N, tri, counts)
    used = eng.last_engine
    want = _run(eng, eng.correlate_dev, dA, dbA, m, dB, dbB, n, W, N, tri, counts)
    assert eng.last_engine == used
    for b in {id(x): x for x in (dA, dbA, dB, dbB)}.values():
        b.free()
    assert (got[0] == want[0]).all(), f"{int((got[0] != want[0]).sum())} r2 cells differ"
    if counts:
        assert (got[1] == want[1]).all(), f"{int((got[1] != want[1]).sum())} counts differ"
    if tri:
        sel = np.triu(np.ones((n, m), dtype=bool), 1)
        assert (got[0][sel] != SENTINEL).all(), "triangle cells left unwritten"
    else:
        assert (got[0] != SENTINEL).all(), "rectangle cells left unwritten"
    return used


@pytest.mark.parametrize("mode", ["reference", "exact"])
def test_noalloc_triangles_match_correlate_dev(mode, engine_factory):
    """12,000 rows (256-tiles), then 3,000 (128-tiles) at 16,622 samples,
    then 2,000 rows of 16 words (the popcount engine under AUTO), each
    identical to qld_correlate_dev."""
    eng = engine_factory(engine="auto", r2_mode=mode)
    for rows, N, counts, engine in ((12000, 16622, False, "mfma_fp4"),
                                    (3000, 16622, True, "mfma_fp4"),
                                    (2000, 1000, True, "popcount")):
        say("noalloc triangle", mode, rows, N)
        A, bc = qld_amd.synth_blocks(20261203 + rows, rows, N)
        assert _case(eng, A, bc, A, bc, N, True, counts) == engine


def test_noalloc_rectangles_match_correlate_dev(engine_factory):
    """Two tables over 100,000 samples: 2,000 x 1,500, then 700 x 300."""
    eng = engine_factory(engine="mfma_fp4", r2_mode="reference")
    N = 100000
    for m, n in ((2000, 1500), (700, 300)):
        say("noalloc rectangle", m, n)
        A, bA, B, bB = qld_amd.synth_regions(20261203 + m, m, n, N)
        assert _case(eng, A, bA, B, bB, N, False, True) == "mfma_fp4"


def test_noalloc_largest_triangle_then_small(engine_factory):
    """A 31,000-row triangle at 16,622 samples (scratch nearly full), then a
    1,100-row one: the second call must not see the first one's tiles."""
    eng = engine_factory(engine="mfma_fp4", r2_mode="exact")
    N = 16622
    for rows, counts in ((31000, False), (1100, True)):
        say("noalloc triangle", rows)
        A, bc = qld_amd.synth_blocks(20261300 + rows, rows, N)
        _case(eng, A, bc, A, bc, N, True, counts)


def test_noalloc_rejects_shapes_beyond_scratch(engine_factory):
    """A triangle whose tiled copy exceeds the 64 MiB scratch is refused on
    the host (QLD_ERR_BUDGET) before anything is launched, and so are
    engines that need a workspace."""
    eng = engine_factory(engine="mfma_fp4", r2_mode="exact")
    N, rows = 16622, 33000
    W = (N + 63) // 64
    dA = eng.alloc(rows * W * 8).fill(0)
    dbc = eng.alloc(rows * 4).fill(0)
    res = eng.alloc(4 * 1024)
    with pytest.raises(qld_amd.QldError, match="scratch"):
        # the output pitch is never used: the call is refused first
        eng.correlate_dev_noalloc(dA, dbc, rows, dA, dbc, rows, W, N, True, res, rows)
    eng.set_fp4_operands("expanded")
    with pytest.raises(qld_amd.QldError, match="workspace"):
        eng.correlate_dev_noalloc(dA, dbc, 2000, dA, dbc, 2000, W, N, True, res, 2000)
    for b in (dA, dbc, res):
        b.free()
