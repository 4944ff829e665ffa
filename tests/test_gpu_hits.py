"""Report selection on the device (qld_correlate_hits, qld_compact_dev).

The reference prints pair (i, j) when r^2 >= r2limit and r^2 < 100, with
j > i when its second window is unset (correlate_IO.c:403-410).  The hits
the device selects -- block by block, streamed to the host -- must be
exactly the cells the host loop over the full r^2 matrix selects, in the
same order and with the same float bits; the report text built from them
must equal the full-matrix writer's text byte for byte.
"""
import os

import numpy as np
import pytest

import oracle
import qld_amd

pytestmark = pytest.mark.gpu

ENGINES = ["popcount", "mfma_fp4", "mfma_i8"]


def host_selection(res, r2limit, upper):
    """The reference's selection loop over a full (n x m) r^2 matrix."""
    n, m = res.shape
    offs = [0]
    cols, vals = [], []
    for i in range(n):
        row = res[i]
        j0 = i + 1 if upper else 0
        sel = np.nonzero((row[j0:] >= np.float32(r2limit)) & (row[j0:] < 100))[0] + j0
        cols.append(sel.astype(np.uint32))
        vals.append(row[sel])
        offs.append(offs[-1] + len(sel))
    return (np.array(offs, np.uint64), np.concatenate(cols) if cols else np.zeros(0, np.uint32),
            np.concatenate(vals) if vals else np.zeros(0, np.float32))


def assert_same_hits(got, exp):
    go, gc, gv = got
    eo, ec, ev = exp
    assert (go == eo).all()
    assert (gc == ec).all()
    assert (gv.view(np.uint32) == ev.view(np.uint32)).all()


def with_special_rows(A, bc, N):
    """Monomorphic, all-ones (NaN r^2) and duplicate rows."""
    A = A.copy()
    bc = bc.copy()
    A[3] = 0
    full = np.zeros(A.shape[1], np.uint64)
    full[: N // 64] = np.uint64(0xFFFFFFFFFFFFFFFF)
    if N % 64:
        full[N // 64] = np.uint64(((1 << (N % 64)) - 1) << (64 - N % 64))
    A[7] = full
    A[11] = A[10]
    bc[3], bc[7], bc[11] = 0, N, bc[10]
    return A, bc


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("block_pairs", [1 << 20, 0])
def test_triangle_hits_equal_full_matrix(engine, block_pairs, engine_factory):
    N = 16622
    A, bc = qld_amd.synth_blocks(11, 1500, N)
    A, bc = with_special_rows(A, bc, N)
    eng = engine_factory(engine=engine)
    full, _ = eng.correlate(A, bc, snp_size=N, triangular=False)
    for lim in (0.2, 0.0, 0.95):
        print(f"[hits] tri engine={engine} blocks={block_pairs} r2limit={lim}", flush=True)
        got = eng.correlate_hits(A, bc, snp_size=N, r2limit=lim, block_pairs=block_pairs)
        assert_same_hits(got, host_selection(full, lim, True))


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("upper", [False, True])
def test_rectangle_hits_equal_full_matrix(engine, upper, engine_factory):
    N = 5000
    A, bcA, B, bcB = qld_amd.synth_regions(12, 700, 1100, N)
    eng = engine_factory(engine=engine)
    full, _ = eng.correlate(A, bcA, B, bcB, snp_size=N)
    print(f"[hits] rect engine={engine} upper={upper}", flush=True)
    got = eng.correlate_hits(A, bcA, B, bcB, snp_size=N, upper=upper, r2limit=0.05,
                             block_pairs=1 << 18)
    assert_same_hits(got, host_selection(full, 0.05, upper))


def test_hits_limits_outside_range(engine_factory):
    N = 3000
    A, bc = qld_amd.synth_blocks(13, 600, N)
    eng = engine_factory(engine="mfma_fp4")
    o, c, v = eng.correlate_hits(A, bc, snp_size=N, r2limit=1.5)
    assert o[-1] == 0 and len(c) == 0
    full, _ = eng.correlate(A, bc, snp_size=N, triangular=False)
    got = eng.correlate_hits(A, bc, snp_size=N, r2limit=-1.0)
    assert_same_hits(got, host_selection(full, -1.0, True))
    assert got[0][-1] == np.count_nonzero(np.triu(full < 100, 1))


def test_compact_dev_matches_host_selection(engine_factory):
    N = 16622
    m = 2000
    A, bc = qld_amd.synth_blocks(14, m, N)
    W = A.shape[1]
    eng = engine_factory(engine="mfma_fp4")
    dA = eng.alloc(A.nbytes).upload(A)
    dbc = eng.alloc(bc.nbytes).upload(bc)
    dres = eng.alloc(m * m * 4)
    eng.correlate_dev(dA, dbc, m, dA, dbc, m, W, N, False, dres, m)
    full = dres.download((m, m), np.float32)
    dcnt = eng.alloc(m * 4)
    doff = eng.alloc((m + 1) * 8)
    cap = m * m
    dcols = eng.alloc(cap * 4)
    dr2 = eng.alloc(cap * 4)
    # rows 500..1499 of the matrix, columns 300.. as one block
    r0, rows, c0 = 500, 1000, 300
    sub = eng.alloc(rows * (m - c0) * 4)
    sub.upload(np.ascontiguousarray(full[r0:r0 + rows, c0:]))
    for upper in (False, True):
        total = eng.compact_dev(sub, m - c0, rows, m - c0, r0, c0, upper, 0.3, dcnt, doff,
                                dcols, dr2, cap)
        offs = doff.download((rows + 1,), np.uint64)
        cols = dcols.download((total,), np.uint32)
        vals = dr2.download((total,), np.float32)
        exp = [], [], [0]
        for i in range(rows):
            row = full[r0 + i]
            j0 = max(c0, r0 + i + 1) if upper else c0
            sel = np.nonzero((row[j0:] >= np.float32(0.3)) & (row[j0:] < 100))[0] + j0
            exp[0].append(sel)
            exp[1].append(row[sel])
            exp[2].append(exp[2][-1] + len(sel))
        assert (offs == np.array(exp[2], np.uint64)).all()
        assert (cols == np.concatenate(exp[0]).astype(np.uint32)).all()
        assert (vals.view(np.uint32) == np.concatenate(exp[1]).view(np.uint32)).all()
    # too small a capacity is reported, not overrun
    with pytest.raises(qld_amd.QldError):
        eng.compact_dev(sub, m - c0, rows, m - c0, r0, c0, False, 0.0, dcnt, doff, dcols, dr2, 10)
    for b in (dA, dbc, dres, dcnt, doff, dcols, dr2, sub):
        b.free()


@pytest.mark.parametrize("posWset2", [0, 1])
def test_streamed_report_equals_full_matrix_writer(posWset2, engine_factory, tmp_path):
    """qld_correlate_report (device selection, streamed) vs the host writer
    over the full matrix (qld_write_results) on the same data."""
    import ctypes as C
    N = 16622
    m = 2500
    A, bc = qld_amd.synth_blocks(15, m, N)
    pos = (np.arange(m, dtype=np.uint32) * 37 + 1000)
    ids = [f"rs{i}" for i in range(m)]
    eng = engine_factory(engine="auto")
    full, _ = eng.correlate(A, bc, snp_size=N, triangular=False)
    ref = oracle.write_results_text(tmp_path / "full.txt", pos, ids, bc, pos, ids, bc, N,
                                    full, 0.2, posWset2)
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    idarr = (C.c_char_p * m)(*[s.encode() for s in ids])
    fp = libc.fopen(str(tmp_path / "dev.txt").encode(), b"w")
    rc = qld_amd.host_lib().qld_correlate_report(
        eng._ctx, fp, A.ctypes.data, pos.ctypes.data, C.cast(idarr, C.c_void_p), bc.ctypes.data,
        m, A.ctypes.data, pos.ctypes.data, C.cast(idarr, C.c_void_p), bc.ctypes.data, m,
        A.shape[1], N, posWset2, C.c_float(0.2))
    libc.fclose(fp)
    assert rc == 0
    assert open(tmp_path / "dev.txt").read() == ref
