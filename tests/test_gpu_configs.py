"""Parity at the benchmark workloads' full shapes, checked against the oracle
on sampled output.  Needs an MI355X.

* 4,096 windows, sizes LogUniform(50, 2000), over a 500,000-SNP x
  16,622-sample chromosome, in one persistent plan (the benchmark's own
  window set), per-window and shared-pair: 48 windows checked in full, the
  largest and the smallest among them.
* Region A 8,000 SNPs x region B 12,000 SNPs over 100,000 samples, the full
  rectangle: 64 rows checked, counts and r^2.
* A server over a 200,000-sample table next to a 16,622-sample one, fed a
  Poisson burst: a 5,000-SNP request equals itself computed alone, bit for
  bit, and 64 of its rows equal the oracle.
The bar is the one of every GPU parity test: counts bit-exact, r^2
bit-identical to the oracle's restatement of the same mode, and within 1e-6
of the FP64 closed form.  The windowing follows the reference's many-task
mode (utilities/split_region:17-39, input_list.c:205-323), the selection of
defined cells writeResults' loop (correlate_IO.c:403-410)."""
import sys
import time

import numpy as np
import pytest

import oracle
import qld_amd

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-6
WAIT_MS = 120000.0


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check_block(got, A_cols, bc_cols, B_rows, bc_rows, N, mode, upper_from=None,
                 got_counts=None):
    """got[r, j] = r^2 of B_rows[r] against A_cols[j]; upper_from[r] = the
    first defined column of row r (triangles), or None (every column)."""
    c = oracle.counts_naive(A_cols, B_rows)
    f = oracle.r2_reference if mode == "reference" else oracle.r2_exact
    want = f(bc_cols, bc_rows, c, N)
    sel = np.ones(got.shape, dtype=bool)
    if upper_from is not None:
        sel = np.arange(got.shape[1])[None, :] >= np.asarray(upper_from)[:, None]
    bad = int((got.view(np.uint32)[sel] != want.view(np.uint32)[sel]).sum())
    assert bad == 0, f"{bad} r2 cells differ"
    if got_counts is not None:
        assert (got_counts[sel] == c[sel]).all(), "counts differ"
    r64 = oracle.r2_fp64(bc_cols, bc_rows, c, N)[sel]
    ok = np.isfinite(r64)
    err = np.abs(got[sel][ok].astype(np.float64) - r64[ok])
    assert err.max(initial=0.0) <= TOL_FP64
    return float(err.max(initial=0.0)), float(err.mean()) if err.size else 0.0


@pytest.mark.parametrize("batch", [qld_amd.BATCH_PERSISTENT, qld_amd.BATCH_SHARED])
def test_config_batched_windows_full_plan(batch, engine_factory):
    N, rows, nwin = 16622, 500000, 4096
    seed = 20261203 + 4
    say("config windows: generating", rows, "SNPs")
    A, bc = qld_amd.synth_blocks(seed, rows, N)
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(50), np.log(2000), nwin))).astype(np.int64)
    starts = rng.integers(0, rows - sizes + 1)
    wins = qld_amd.window_layout(starts, sizes)
    total = int((sizes * sizes).sum())
    e = engine_factory(engine="auto", r2_mode="exact")
    t = e.table(A, bc, N)
    out = e.alloc(total * 4)
    out.fill(0x7FC00000)
    plan = e.plan_windows(t, wins, out, batch)
    st = plan.stats()
    assert st["pairs"] == int((sizes * (sizes - 1) // 2).sum())
    say("config windows: plan", batch, st)
    plan.run()
    e.sync()
    pick = {int(np.argmax(sizes)), int(np.argmin(sizes))}
    pick |= set(int(k) for k in np.random.default_rng(7).choice(nwin, 46, replace=False))
    worst = 0.0
    for k in sorted(pick):
        s, n, off = int(wins[k]["start"]), int(wins[k]["size"]), int(wins[k]["out_offset"])
        got = out.download_at(off * 4, (n, n), np.float32)
        sub = np.ascontiguousarray(A[s:s + n])
        mx, _ = _check_block(got, sub, bc[s:s + n], sub, bc[s:s + n], N, "exact",
                             upper_from=np.arange(n) + 1)
        worst = max(worst, mx)
    say("config windows: checked", len(pick), "windows, max |r2 - fp64|", worst)
    plan.free()
    t.free()
    out.free()


@pytest.mark.parametrize("mode", ["reference", "exact"])
def test_config_rectangle_100k_full(mode, engine_factory):
    N, m, n = 100000, 8000, 12000
    say("config rectangle", mode)
    A, bA, B, bB = qld_amd.synth_regions(20261203 + 3, m, n, N)
    W = A.shape[1]
    e = engine_factory(engine="auto", r2_mode=mode)
    dA, dbA = e.alloc(A.nbytes).upload(A), e.alloc(bA.nbytes).upload(bA)
    dB, dbB = e.alloc(B.nbytes).upload(B), e.alloc(bB.nbytes).upload(bB)
    dres, dcnt = e.alloc(m * n * 4), e.alloc(m * n * 4)
    e.correlate_dev(dA, dbA, m, dB, dbB, n, W, N, False, dres, m, dcnt, m)
    e.sync()
    assert e.last_engine == "mfma_fp4"
    rows = np.sort(np.random.default_rng(3).choice(n, 64, replace=False))
    rows[0], rows[-1] = 0, n - 1
    got = np.stack([dres.download_at(int(r) * m * 4, m, np.float32) for r in rows])
    gotc = np.stack([dcnt.download_at(int(r) * m * 4, m, np.uint32) for r in rows])
    mx, mean = _check_block(got, A, bA, np.ascontiguousarray(B[rows]), bB[rows], N, mode,
                            got_counts=gotc)
    say("config rectangle", mode, "max |r2 - fp64|", mx, "mean", mean)
    for b in (dA, dbA, dB, dbB, dres, dcnt):
        b.free()


def test_config_server_200k_samples_poisson_burst(engine_factory):
    e = engine_factory(engine="auto", r2_mode="reference")
    say("config server: tables")
    A2, b2 = qld_amd.synth_blocks(20261203 + 53, 20000, 200000)
    A1, b1 = qld_amd.synth_blocks(20261203 + 50, 30000, 16622)
    t2, t1 = e.table(A2, b2, 200000), e.table(A1, b1, 16622)
    tabs = [(t2, 20000), (t1, 30000)]
    big_start, big = 7000, 5000
    srv = qld_amd.Server(e, out_budget=4 << 30, max_batch=64, streams=3)
    # alone
    rid = srv.submit(t2, big_start, big)
    srv.wait(rid, WAIT_MS)
    alone = srv.fetch(rid, big)
    # a Poisson burst at 5,000 requests/s with the same request in the middle
    rng = np.random.default_rng(5)
    ids, mine = [], None
    due = time.perf_counter()
    for k in range(300):
        due += rng.exponential(1.0 / 5000.0)
        while time.perf_counter() < due:
            pass
        if k == 150:
            mine = srv.submit(t2, big_start, big)
            continue
        tab, nrows = tabs[int(rng.integers(2))]
        size = int(np.rint(np.exp(rng.uniform(np.log(20), np.log(5000)))))
        ids.append(srv.submit(tab, int(rng.integers(0, nrows - size + 1)), size,
                              qld_amd.REQ_DISCARD))
    srv.wait(mine, WAIT_MS)
    mixed = srv.fetch(mine, big)
    for i in ids:
        srv.wait(i, WAIT_MS)
        srv.release(i)
    st = srv.stats()
    say("config server: stats", st)
    srv.close()
    sel = np.triu(np.ones((big, big), dtype=bool), 1)
    assert (alone.view(np.uint32)[sel] == mixed.view(np.uint32)[sel]).all(), \
        "alone and in the mix differ"
    rows = np.sort(np.random.default_rng(9).choice(big, 64, replace=False))
    win = np.ascontiguousarray(A2[big_start:big_start + big])
    mx, _ = _check_block(alone[rows], win, b2[big_start:big_start + big],
                         np.ascontiguousarray(win[rows]), b2[big_start + rows], 200000,
                         "reference", upper_from=rows + 1)
    say("config server: max |r2 - fp64|", mx)
    t1.free()
    t2.free()
