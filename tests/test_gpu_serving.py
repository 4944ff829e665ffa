"""Serving loop: results against the oracle, alone == in the mix (bit for
bit), admission by output budget, copy-out mode.  Needs an MI355X."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle
import qld_amd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def tables(engine_factory):
    e = engine_factory(engine="auto", r2_mode="reference")
    A16, b16 = qld_amd.synth_blocks(161, 3000, 16622)
    A100, b100 = qld_amd.synth_blocks(100, 700, 100000)
    # one table of each fp4 operand form, so that batches mix the two
    # launch groups (packed rows into registers / expanded copy via LDS)
    t16 = e.table(A16, b16, 16622)
    e.set_fp4_operands("expanded")
    t100 = e.table(A100, b100, 100000)
    e.set_fp4_operands("direct")
    yield e, [(t16, A16, b16, 16622), (t100, A100, b100, 100000)]
    t16.free()
    t100.free()


def _want(A, bc, N, start, size):
    sub = np.ascontiguousarray(A[start:start + size])
    return oracle.r2_reference(bc[start:start + size], bc[start:start + size],
                               oracle.counts_naive(sub), N)


WAIT_MS = 60000.0     # every wait is bounded: a stuck server fails the test


def _tri(x):
    n = x.shape[0]
    return x[np.triu(np.ones((n, n), dtype=bool), 1)]


def _requests(rng, tabs, k, lo, hi):
    out = []
    for _ in range(k):
        ti = int(rng.integers(len(tabs)))
        rows = tabs[ti][1].shape[0]
        size = int(np.clip(np.exp(rng.uniform(np.log(lo), np.log(hi))), 1, rows))
        start = int(rng.integers(0, rows - size + 1))
        out.append((ti, start, size))
    return out


def test_server_matches_oracle(tables):
    e, tabs = tables
    srv = qld_amd.Server(e, out_budget=256 << 20, max_batch=16, streams=3)
    reqs = _requests(np.random.default_rng(0), tabs, 40, 2, 500)
    ids = [srv.submit(tabs[t][0], s, n) for t, s, n in reqs]
    for rid, (t, s, n) in zip(ids, reqs):
        lat = srv.wait(rid, WAIT_MS)
        assert lat > 0
        got = srv.fetch(rid, n)
        want = _want(tabs[t][1], tabs[t][2], tabs[t][3], s, n)
        assert (_tri(got).view(np.uint32) == _tri(want).view(np.uint32)).all(), (t, s, n)
    st = srv.stats()
    say("server stats", st)
    assert st["completed"] == len(reqs) and st["status"] == 0
    srv.close()


def test_alone_equals_in_mix(tables):
    e, tabs = tables
    t, A, bc, N = tabs[0]
    probe = (123, 450)
    srv = qld_amd.Server(e, out_budget=256 << 20, max_batch=8, streams=2)
    rid = srv.submit(t, *probe)
    srv.wait(rid, WAIT_MS)
    alone = srv.fetch(rid, probe[1])
    # the same request in the middle of a burst of others on both tables
    mix = _requests(np.random.default_rng(1), tabs, 30, 20, 600)
    ids = [srv.submit(tabs[ti][0], s, n, qld_amd.REQ_DISCARD) for ti, s, n in mix[:15]]
    rid = srv.submit(t, *probe)
    ids += [srv.submit(tabs[ti][0], s, n, qld_amd.REQ_DISCARD) for ti, s, n in mix[15:]]
    srv.wait(rid, WAIT_MS)
    in_mix = srv.fetch(rid, probe[1])
    for i in ids:
        srv.wait(i, WAIT_MS)
        srv.release(i)
    assert (_tri(alone).view(np.uint32) == _tri(in_mix).view(np.uint32)).all()
    assert (_tri(alone).view(np.uint32) == _tri(_want(A, bc, N, *probe)).view(np.uint32)).all()
    st = srv.stats()
    say("mix stats", st)
    assert st["batches"] >= 2
    srv.close()


def test_budget_admission(tables):
    e, tabs = tables
    t, A, bc, N = tabs[0]
    size = 400                                   # 640 KB of output each
    srv = qld_amd.Server(e, out_budget=2 * size * size * 4, max_batch=4, streams=2)
    with pytest.raises(qld_amd.QldError) as x:
        srv.submit(t, 0, 1000)                   # can never fit
    assert x.value.code == qld_amd.QLD_ERR_BUDGET
    starts = list(range(0, 2400, 200))
    ids = [srv.submit(t, s, size) for s in starts]   # at most 2 in flight
    for rid, s in zip(ids, starts):
        srv.wait(rid, WAIT_MS)
        got = srv.fetch(rid, size)
        assert (_tri(got).view(np.uint32) == _tri(_want(A, bc, N, s, size)).view(np.uint32)).all()
    srv.close()


def test_copy_mode_pinned(tables):
    e, tabs = tables
    t, A, bc, N = tabs[1]
    size = 300
    p = C.c_void_p()
    assert qld_amd.hip_lib().qld_host_alloc(e._ctx, size * size * 4, C.byref(p)) == 0
    srv = qld_amd.Server(e, out_budget=64 << 20)
    rid = srv.submit(t, 50, size, qld_amd.REQ_COPY, p)
    srv.wait(rid, WAIT_MS)
    got = np.ctypeslib.as_array((C.c_float * (size * size)).from_address(p.value)).reshape(size, size).copy()
    assert (_tri(got).view(np.uint32) == _tri(_want(A, bc, N, 50, size)).view(np.uint32)).all()
    srv.release(rid)
    srv.close()
    qld_amd.hip_lib().qld_host_free(e._ctx, p)


@pytest.mark.parametrize("chunks", [0, 1, 5, 64])
def test_split_batches_match_oracle(chunks, engine_factory):
    """Batches on a direct table of long rows (100,000 samples, 391 word
    groups) split their tiles' k loops by shape (0), never (1) or into a
    forced chunk count (5, 64: more chunks than some tiles' batches need):
    every result against the oracle, alone and batched."""
    e = engine_factory(engine="auto", r2_mode="reference")
    e.set_ksplit(chunks)
    N = 100000
    A, bc = qld_amd.synth_blocks(1000 + chunks, 900, N)
    t = e.table(A, bc, N)
    srv = qld_amd.Server(e, out_budget=256 << 20, max_batch=6, streams=2)
    reqs = [(s, n) for (_, s, n) in _requests(np.random.default_rng(chunks), [(t, A)], 16, 2, 400)]
    say("split batches", chunks, len(reqs))
    ids = [srv.submit(t, s, n) for s, n in reqs]
    for rid, (s, n) in zip(ids, reqs):
        srv.wait(rid, WAIT_MS)
        got = srv.fetch(rid, n)
        assert (_tri(got).view(np.uint32) == _tri(_want(A, bc, N, s, n)).view(np.uint32)).all(), (s, n)
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["status"] == 0
    srv.close()
    t.free()
