"""Batched windows on a resident table: plain (one launch per window) and
persistent (one launch per tile shape) plans, direct and graph-replayed,
against the oracle window by window.  Needs an MI355X."""
import sys

import numpy as np
import pytest

import oracle
import qld_amd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _windows(rng, rows, k, lo, hi):
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), k)).round().astype(np.int64)
    sizes = np.clip(sizes, 1, rows)
    starts = rng.integers(0, rows - sizes + 1)
    return qld_amd.window_layout(starts, sizes)


def _check_windows(out, wins, A, bc, N, mode="reference"):
    f = oracle.r2_reference if mode == "reference" else oracle.r2_exact
    for w in wins:
        s, n, off = int(w["start"]), int(w["size"]), int(w["out_offset"])
        if n < 2:
            continue
        got = out[off:off + n * n].reshape(n, n)
        sub = np.ascontiguousarray(A[s:s + n])
        want = f(bc[s:s + n], bc[s:s + n], oracle.counts_naive(sub), N)
        sel = np.triu(np.ones((n, n), dtype=bool), 1)
        bad = int((got.view(np.uint32)[sel] != want.view(np.uint32)[sel]).sum())
        assert bad == 0, f"window start={s} size={n}: {bad} cells differ"


@pytest.mark.parametrize("engine", ["mfma_fp4", "mfma_fp4x", "mfma_i8"])
def test_window_plan_sample_count_above_row_bits(engine, engine_factory):
    """A resident table whose N (the header's sample count, 22,554) exceeds
    its rows' bits (5,638 listed samples), as qLD-compute -inputList builds
    with a sample list: windows against the oracle with that N."""
    N, bits, rows = 22554, 5638, 900
    A, bc = qld_amd.synth_blocks(4343, rows, bits)
    wins = qld_amd.window_layout(np.array([0, 37, 300, 512], np.int64),
                                 np.array([300, 129, 388, 257], np.int64))
    total = int((wins["size"].astype(np.int64) ** 2).sum())
    e = engine_factory(engine=engine, r2_mode="reference")
    t = e.table(A, bc, N)
    out = e.alloc(total * 4)
    plan = e.plan_windows(t, wins, out, qld_amd.BATCH_PERSISTENT)
    say("plan N above row bits", engine)
    out.fill(0x7FC00000)
    plan.run()
    e.sync()
    _check_windows(out.download(total, np.float32), wins, A, bc, N)
    plan.free()
    t.free()
    out.free()


@pytest.mark.parametrize("engine", ["mfma_fp4", "mfma_fp4x", "mfma_i8"])
@pytest.mark.parametrize("mode", [qld_amd.BATCH_PERSISTENT, qld_amd.BATCH_PLAIN])
@pytest.mark.parametrize("cutoff", [300, 0, 1 << 30])
def test_window_plans(engine, mode, cutoff, engine_factory):
    N, rows = 16622, 4000
    A, bc = qld_amd.synth_blocks(4242, rows, N)
    rng = np.random.default_rng(cutoff % 97 + mode)
    wins = _windows(rng, rows, 40, 2, 700)
    wins[0]["size"] = 1                     # a window without pairs
    wins = qld_amd.window_layout(wins["start"], wins["size"])
    total = int((wins["size"].astype(np.int64) ** 2).sum())
    e = engine_factory(engine=engine, r2_mode="reference")
    e.set_tile_cutoff(cutoff)
    t = e.table(A, bc, N)
    out = e.alloc(total * 4)
    plan = e.plan_windows(t, wins, out, mode)
    st = plan.stats()
    say("plan", engine, mode, cutoff, st)
    assert st["pairs"] == int(sum(int(s) * (int(s) - 1) // 2 for s in wins["size"]))
    for use_graph in (True, False, True):
        out.fill(0x7FC00000)                # NaN everywhere before each run
        plan.use_graph(use_graph)
        plan.run()
        e.sync()
        _check_windows(out.download(total, np.float32), wins, A, bc, N)
    plan.free()
    t.free()
    out.free()


@pytest.mark.parametrize("sched", ["tiles", "waves"])
@pytest.mark.parametrize("cutoff", [300, 0, 1 << 30])
def test_window_plans_both_schedules(sched, cutoff, engine_factory):
    """Persistent fp4 plans with whole tiles per workgroup and with wave
    blocks claimed per wave, 128- and 256-tiles, every window size from 1
    to 1,100 SNPs in the mix (ragged edges and diagonals of both wave block
    shapes), in both r^2 modes."""
    N, rows = 16622, 6000
    A, bc = qld_amd.synth_blocks(4244, rows, N)
    rng = np.random.default_rng(cutoff % 89 + (sched == "tiles"))
    wins = _windows(rng, rows, 48, 2, 1100)
    wins = qld_amd.window_layout(np.concatenate([wins["start"], [0, 5, 700, 1000]]),
                                 np.concatenate([wins["size"], [1, 256, 512, 1024]]))
    total = int((wins["size"].astype(np.int64) ** 2).sum())
    for mode in ("reference", "exact"):
        e = engine_factory(engine="mfma_fp4", r2_mode=mode)
        e.set_tile_cutoff(cutoff)
        e.set_plan_schedule(sched)
        t = e.table(A, bc, N)
        out = e.alloc(total * 4)
        plan = e.plan_windows(t, wins, out, qld_amd.BATCH_PERSISTENT)
        say("plan", sched, cutoff, mode, plan.stats())
        for use_graph in (True, False):
            out.fill(0x7FC00000)
            plan.use_graph(use_graph)
            plan.run()
            e.sync()
            _check_windows(out.download(total, np.float32), wins, A, bc, N, mode)
        plan.free()
        t.free()
        out.free()


def test_window_plan_without_graph_capture(engine_factory):
    """A plan built with graph capture switched off runs by direct launches
    (and refuses graph replay) with the same results."""
    N, rows = 16622, 3000
    A, bc = qld_amd.synth_blocks(4245, rows, N)
    wins = _windows(np.random.default_rng(11), rows, 24, 2, 900)
    total = int((wins["size"].astype(np.int64) ** 2).sum())
    e = engine_factory(engine="mfma_fp4", r2_mode="reference")
    e.set_plan_graphs(False)
    t = e.table(A, bc, N)
    out = e.alloc(total * 4)
    plan = e.plan_windows(t, wins, out, qld_amd.BATCH_PERSISTENT)
    with pytest.raises(qld_amd.QldError):
        plan.use_graph(True)
    out.fill(0x7FC00000)
    plan.run()
    e.sync()
    _check_windows(out.download(total, np.float32), wins, A, bc, N)
    plan.free()
    t.free()
    out.free()


@pytest.mark.parametrize("engine", ["mfma_fp4", "mfma_fp4x"])
def test_table_from_device_rows(engine, engine_factory):
    """A table built from device buffers (qld_table_create_dev) gives the
    same windows as one built from host rows; rows not a multiple of 32 and
    windows reaching the last row exercise the tiled copy's edges."""
    N, rows = 100000, 1037
    A, bc = qld_amd.synth_blocks(777, rows, N)
    rng = np.random.default_rng(5)
    wins = _windows(rng, rows, 12, 40, 600)
    wins = qld_amd.window_layout(np.append(wins["start"], rows - 300),
                                 np.append(wins["size"], 300))
    total = int((wins["size"].astype(np.int64) ** 2).sum())
    e = engine_factory(engine=engine, r2_mode="reference")
    dA = e.alloc(A.nbytes).upload(A)
    dbc = e.alloc(bc.nbytes).upload(bc)
    t = qld_amd.Table.from_device(e, dA, dbc, rows, A.shape[1], N)
    dA.fill(0)                              # the table holds its own copy
    assert t.nbytes > 0
    out = e.alloc(total * 4)
    out.fill(0x7FC00000)
    plan = e.plan_windows(t, wins, out)
    plan.run()
    e.sync()
    _check_windows(out.download(total, np.float32), wins, A, bc, N)
    for b in (plan, t, out, dA, dbc):
        b.free()


def test_window_plan_exact_mode_and_samples_100k(engine_factory):
    N, rows = 100000, 900
    A, bc = qld_amd.synth_blocks(99, rows, N)
    wins = _windows(np.random.default_rng(1), rows, 12, 20, 600)
    total = int((wins["size"].astype(np.int64) ** 2).sum())
    e = engine_factory(engine="auto", r2_mode="exact")
    e.set_tile_cutoff(400)
    t = e.table(A, bc, N)
    out = e.alloc(total * 4)
    plan = e.plan_windows(t, wins, out, qld_amd.BATCH_PERSISTENT)
    plan.run()
    e.sync()
    _check_windows(out.download(total, np.float32), wins, A, bc, N, mode="exact")
    plan.free()
    t.free()
    out.free()


def test_window_validation(engine_factory):
    e = engine_factory(engine="auto")
    A, bc = qld_amd.synth_blocks(5, 100, 1000)
    t = e.table(A, bc, 1000)
    out = e.alloc(4 * 100 * 100)
    bad = qld_amd.window_layout([90], [20])          # runs past the table
    with pytest.raises(qld_amd.QldError) as x:
        e.plan_windows(t, bad, out)
    assert x.value.code == qld_amd.QLD_ERR_INVALID
    t.free()
    out.free()


def test_autotuner_picks_a_candidate(engine_factory):
    N, rows = 16622, 6000
    A, bc = qld_amd.synth_blocks(7, rows, N)
    wins = _windows(np.random.default_rng(2), rows, 64, 50, 2000)
    total = int((wins["size"].astype(np.int64) ** 2).sum())
    e = engine_factory(engine="auto")
    t = e.table(A, bc, N)
    out = e.alloc(total * 4)
    best, ms = e.autotune_windows(t, wins, out, reps=2)
    say("autotune best cutoff", best, "ms", ms)
    assert best in (0, 256, 384, 512, 768, 1024, 1536, 2048, 1 << 30)
    assert e.tile_cutoff == best and ms > 0
    # the tuned plan still computes the right answer
    plan = e.plan_windows(t, wins, out)
    plan.run()
    e.sync()
    _check_windows(out.download(total, np.float32), wins[:16], A, bc, N)
    plan.free()
    t.free()
    out.free()


@pytest.mark.parametrize("engine", ["mfma_fp4", "mfma_fp4x", "mfma_i8", "mfma_fp4:tiles"])
@pytest.mark.parametrize("mode", [qld_amd.BATCH_PERSISTENT, qld_amd.BATCH_PLAIN])
def test_pair_window_plans_mix_triangles_and_rectangles(engine, mode, engine_factory):
    """qld_plan_pairs: triangles and rectangles (overlapping, disjoint,
    empty, 1-row) in one plan, each against the oracle (fp4 with wave
    blocks claimed per wave, the default, and with whole tiles)."""
    engine, _, sched = engine.partition(":")
    N, rows = 16622, 3000
    A, bc = qld_amd.synth_blocks(777, rows, N)
    rng = np.random.default_rng(mode * 7 + (engine == "mfma_i8"))
    spec = []
    for k in range(30):
        a = int(rng.integers(1, 600))
        b = int(rng.integers(1, 600))
        a0 = int(rng.integers(0, rows - a + 1))
        b0 = int(rng.integers(0, rows - b + 1))
        spec.append((a0, a, a0, a, 1) if k % 3 == 0 else (a0, a, b0, b, 0))
    spec += [(10, 0, 20, 5, 0), (100, 1, 200, 1, 0), (50, 1, 50, 1, 1)]
    w = np.zeros(len(spec), dtype=qld_amd.PAIR_WINDOW_DTYPE)
    off = 0
    for k, (a0, a, b0, b, tri) in enumerate(spec):
        w[k] = (a0, a, b0, b, off, tri, 0)
        off += a * b
    e = engine_factory(engine=engine, r2_mode="reference")
    e.set_tile_cutoff(400)
    if sched:
        e.set_plan_schedule(sched)
    t = e.table(A, bc, N)
    out = e.alloc(max(off, 1) * 4)
    plan = e.plan_windows(t, w, out, mode)
    want_pairs = sum(a * (a - 1) // 2 if tri else a * b for a0, a, b0, b, tri in spec)
    assert plan.stats()["pairs"] == want_pairs
    for use_graph in (True, False):
        out.fill(0x7FC00000)
        plan.use_graph(use_graph)
        plan.run()
        e.sync()
        got_all = out.download(max(off, 1), np.float32)
        for k, (a0, a, b0, b, tri) in enumerate(spec):
            if a == 0 or b == 0 or (tri and a < 2):
                continue
            o = int(w[k]["out_offset"])
            got = got_all[o:o + a * b].reshape(b, a)
            want = oracle.r2_reference(bc[a0:a0 + a], bc[b0:b0 + b],
                                       oracle.counts_naive(A[a0:a0 + a], A[b0:b0 + b]), N)
            sel = np.triu(np.ones((b, a), dtype=bool), 1) if tri else np.ones((b, a), bool)
            bad = int((got.view(np.uint32)[sel] != want.view(np.uint32)[sel]).sum())
            say("pair window", k, (a0, a, b0, b, tri), "bad", bad)
            assert bad == 0
    plan.free()
    t.free()
    out.free()


def test_pair_window_validation(engine_factory):
    e = engine_factory(engine="auto")
    A, bc = qld_amd.synth_blocks(5, 100, 1000)
    t = e.table(A, bc, 1000)
    out = e.alloc(4 * 100 * 100)
    for bad in [(0, 10, 95, 10, 0), (0, 10, 5, 10, 1), (-1, 5, 0, 5, 0)]:
        w = np.zeros(1, dtype=qld_amd.PAIR_WINDOW_DTYPE)
        w[0] = (bad[0], bad[1], bad[2], bad[3], 0, bad[4], 0)
        with pytest.raises(qld_amd.QldError) as x:
            e.plan_windows(t, w, out)
        assert x.value.code == qld_amd.QLD_ERR_INVALID
    t.free()
    out.free()


@pytest.mark.parametrize("r2_mode", ["reference", "exact"])
def test_shared_pair_plans(r2_mode, engine_factory):
    """QLD_BATCH_SHARED: overlapping windows -- random, nested, duplicated,
    sizes 0 to 3, one reaching the table's last row (5,003 rows: no multiple
    of the 128 x 64 blocks) -- each pair counted once and written to every
    window holding it: every window against the oracle and bit-identical to
    the per-window plan, graph-replayed and by direct launches."""
    N, rows = 16622, 5003
    A, bc = qld_amd.synth_blocks(4246, rows, N)
    rng = np.random.default_rng(31 + (r2_mode == "exact"))
    wins = _windows(rng, rows, 40, 2, 1500)
    extra_s = [0, 0, 0, 64, 127, 128, 1000, 1000, 1100, rows - 700, rows - 2, rows - 1, 2500]
    extra_n = [0, 1, 2, 3, 129, 64, 800, 800, 300, 700, 2, 1, 1500]
    wins = qld_amd.window_layout(np.concatenate([wins["start"], extra_s]),
                                 np.concatenate([wins["size"], extra_n]))
    total = max(int((wins["size"].astype(np.int64) ** 2).sum()), 1)
    e = engine_factory(engine="mfma_fp4", r2_mode=r2_mode)
    t = e.table(A, bc, N)
    out = e.alloc(total * 4)
    ref_out = e.alloc(total * 4)
    per_window = e.plan_windows(t, wins, ref_out, qld_amd.BATCH_PERSISTENT)
    ref_out.fill(0x7FC00000)
    per_window.run()
    e.sync()
    want_all = ref_out.download(total, np.float32)
    shared = e.plan_windows(t, wins, out, qld_amd.BATCH_SHARED)
    st = shared.stats()
    say("shared plan", r2_mode, st, "per-window", per_window.stats())
    assert st["pairs"] == per_window.stats()["pairs"] and st["launches"] == 1
    for use_graph in (True, False):
        out.fill(0x7FC00000)
        shared.use_graph(use_graph)
        shared.run()
        e.sync()
        got_all = out.download(total, np.float32)
        for w in wins:
            s, n, off = int(w["start"]), int(w["size"]), int(w["out_offset"])
            if n < 2:
                continue
            sel = np.triu(np.ones((n, n), dtype=bool), 1)
            g = got_all[off:off + n * n].reshape(n, n)[sel].view(np.uint32)
            r = want_all[off:off + n * n].reshape(n, n)[sel].view(np.uint32)
            assert (g == r).all(), f"window start={s} size={n} differs from the per-window plan"
        _check_windows(got_all, wins, A, bc, N, r2_mode)
    for b in (shared, per_window, t, out, ref_out):
        b.free()


def test_shared_pair_plan_validation(engine_factory):
    """Shared-pair plans take single-region windows of a direct fp4 table."""
    A, bc = qld_amd.synth_blocks(5, 300, 1000)
    e = engine_factory(engine="mfma_fp4")
    t = e.table(A, bc, 1000)
    out = e.alloc(4 * 300 * 300)
    w = np.zeros(1, dtype=qld_amd.PAIR_WINDOW_DTYPE)
    w[0] = (0, 10, 20, 10, 0, 0, 0)                   # a rectangle
    with pytest.raises(qld_amd.QldError) as x:
        e.plan_windows(t, w, out, qld_amd.BATCH_SHARED)
    assert x.value.code == qld_amd.QLD_ERR_INVALID
    # with the shared form preferred, a persistent plan of pair windows that
    # include a rectangle keeps the per-window form, and stays correct
    e.set_plan_shared(True)
    assert e.plan_shared
    w2 = np.zeros(2, dtype=qld_amd.PAIR_WINDOW_DTYPE)
    w2[0] = (0, 100, 150, 120, 0, 0, 0)
    w2[1] = (40, 200, 40, 200, 100 * 120, 1, 0)
    out.fill(0x7FC00000)
    plan = e.plan_windows(t, w2, out, qld_amd.BATCH_PERSISTENT)
    plan.run()
    e.sync()
    got = out.download(100 * 120 + 200 * 200, np.float32)
    want = oracle.r2_reference(bc[0:100], bc[150:270], oracle.counts_naive(A[0:100], A[150:270]),
                               1000)
    assert (got[:12000].reshape(120, 100).view(np.uint32) == want.view(np.uint32)).all()
    plan.free()
    t.free()
    for engine in ("mfma_fp4x", "mfma_i8"):            # expanded operands
        e2 = engine_factory(engine=engine)
        t2 = e2.table(A, bc, 1000)
        with pytest.raises(qld_amd.QldError) as x:
            e2.plan_windows(t2, qld_amd.window_layout([0], [100]), out,
                            qld_amd.BATCH_SHARED)
        assert x.value.code == qld_amd.QLD_ERR_INVALID
        t2.free()
    out.free()
