"""Benchmark of the LD hot path on one MI355X.

Default workload: one region of 30,000 SNPs x 16,622 haploid samples
(seeded haplotype-block Markov data), all-pairs r^2 over the lower triangle
(449,985,000 pairs), packed table resident in HBM.  A step = one full pass
(contraction + fused r^2 epilogue, FP32 r^2 tiles written to HBM).

Prints ONE JSON line (rank 0):
  value         pairs / second of the whole step, Gpairs/s
  roofline      dominant kernel's achieved rate vs the MI355X peak of the
                unit it is bound by, measured with HIP events on the engine's
                stream
  cpu_baseline  the oracle's C/OpenMP restatement of the reference's BLIS CPU
                path (blocked AND+POPCNT + get_pairwise_ld_score epilogue),
                timed on a bounded sample of the same workload on this host
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = ("SNP-pair r²/sec (Gpairs/s) at 16,622 samples; batched vs plain GPU, "
          "p99 latency")

# MI355X peaks used as roofline denominators.
PEAK = {
    "valu_int32": 39.3e12,   # 256 CU x 64 lanes x 2.4 GHz (int32 VALU ops/s)
    "mfma_fp4": 10.1e15,     # dense fp4 MFMA
    "mfma_i8": 5.0e15,       # dense int8 MFMA
    "hbm": 8.0e12,
}
DTYPE = {"popcount": "u64", "mfma_i8": "int8", "mfma_fp4": "fp4"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(workload, r2_mode):
    """HBM bytes per launch of the contraction kernel for this workload, read
    from the committed rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE in
    KiB per dispatch, each scaled by what it reads for a copy kernel of known
    size); None when no pass covers it."""
    key = f"workload={workload} mode={r2_mode} "
    try:
        for path in sorted(os.listdir(os.path.join(ROOT, "profiles"))):
            with open(os.path.join(ROOT, "profiles", path)) as f:
                for line in f:
                    if line.startswith("traffic_bytes ") and key in line + " ":
                        kv = dict(t.split("=", 1) for t in line.split()[1:])
                        return {"bytes": int(kv["fetch"]) + int(kv["write"]),
                                "fetch": int(kv["fetch"]), "write": int(kv["write"]),
                                "source": f"profiles/{path}"}
    except OSError:
        pass
    return None


def roofline(engine, pairs, words, samples, kernel_s, traffic=None, rows=None, out_pairs=None):
    """pairs: the pairs the kernel computes (the flop count); out_pairs: the
    r2 values it writes when that differs (shared-pair plans)."""
    if engine == "popcount":
        ops = pairs * 4.0 * words          # 2x v_and_b32 + 2x v_bcnt per word pair
        ach = ops / kernel_s
        return {"bound": "valu", "achieved": ach / 1e12, "peak": PEAK["valu_int32"] / 1e12,
                "unit": "Tops/s (int32 VALU)", "frac": ach / PEAK["valu_int32"],
                "traffic": None, "kernel": "ld_popcount_kernel",
                "ops_per_pair": 4 * words, "kernel_ms": kernel_s * 1e3}
    flops = pairs * 2.0 * samples
    ach = flops / kernel_s
    pk = PEAK[engine]
    out = {"bound": "mfma", "achieved": ach / 1e12, "peak": pk / 1e12,
           "unit": "TFLOP/s", "frac": ach / pk,
           "traffic": traffic["bytes"] if traffic else None,
           "flops_per_pair": 2 * samples, "kernel_ms": kernel_s * 1e3}
    if traffic:
        out["traffic_detail"] = traffic
        if rows:
            # each operand row read once (fp4: 4 bits per sample, padded to the
            # 64-sample word) + 4 B of r2 written per pair
            opb = {"mfma_i8": 64}.get(engine, 8)   # fp4 reads the packed rows
            out["algorithmic_bytes"] = rows * words * opb + (out_pairs or pairs) * 4
    return out


def host_cpu():
    """Physical cores (lscpu: sockets x cores per socket), the CPUs this
    process may run on, the thread count the box sets, and the CPU model."""
    import subprocess
    info = {"model": "unknown", "physical_cores": None,
            "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["model"] = kv.get("Model name", "unknown")
        info["physical_cores"] = int(kv["Socket(s)"]) * int(kv["Core(s) per socket"])
    except (OSError, KeyError, ValueError, subprocess.SubprocessError):
        pass
    env = os.environ.get("OMP_NUM_THREADS")
    info["omp_num_threads"] = int(env) if env and env.isdigit() else None
    return info


def cpu_baseline(A, bc, N, budget_s=10.0):
    """The oracle's C restatement of the reference's BLIS CPU path (blocked
    AND+popcount counts, get_pairwise_ld_score's r^2) on the first S SNPs of
    the workload (triangle), on every physical core this process may use and
    on one core; with the calibration against the reference's own BLIS build
    (oracle/calibrate_blis.py, run where the reference can be built)."""
    import oracle
    cpu = host_cpu()
    threads = min(x for x in (cpu["physical_cores"], cpu["affinity_cpus"],
                              cpu["omp_num_threads"]) if x)

    def run(S, T):
        sub = np.ascontiguousarray(A[:S])
        t0 = time.perf_counter()
        c = oracle.counts_blocked(sub, tri=True, nthreads=T)
        oracle.r2_reference(bc[:S], bc[:S], c, N, tri=True)
        return time.perf_counter() - t0

    def sample(T, budget):
        S = min(3000, A.shape[0])
        t = run(S, T)
        S2 = int(min(A.shape[0], S * (budget / max(t, 1e-3)) ** 0.5))
        if S2 > S:
            S = S2
            t = run(S, T)
        return S, t

    S, t = sample(threads, budget_s)
    S1, t1 = sample(1, budget_s / 2)
    pairs, pairs1 = S * (S - 1) / 2, S1 * (S1 - 1) / 2
    calib = None
    try:
        with open(os.path.join(ROOT, "oracle", "blis_calibration.json")) as f:
            cal = json.load(f)
        calib = {"measured_on": cal["host"],
                 "restatement_vs_reference_blis": {
                     f"{c['snps']}_snps_{k}": c[k]["restatement_speedup"]
                     for c in cal["configs"] for k in c if k.startswith("threads_")},
                 "reference_blis_counts_r2_s": {
                     f"{c['snps']}_snps_{k}": c[k]["reference_counts_r2_s"]
                     for c in cal["configs"] for k in c if k.startswith("threads_")}}
    except (OSError, KeyError, ValueError):
        pass
    return {"value": pairs / t / 1e9, "unit": "Gpairs/s", "cores": threads, "kind": "port",
            "physical_cores": cpu["physical_cores"], "cpu_model": cpu["model"],
            "popcount_bits": oracle.counts_blocked_isa(),
            "sample": f"first {S} SNPs x {N} samples, lower triangle "
                      f"({int(pairs)} pairs), counts + r^2, {t:.2f} s on {threads} threads",
            "one_core": {"value": pairs1 / t1 / 1e9, "unit": "Gpairs/s",
                         "sample": f"first {S1} SNPs, {int(pairs1)} pairs, {t1:.2f} s"},
            "calibration_vs_blis": calib}


def r2_error_vs_fp64(eng, dA, dbc, m, W, N, rows):
    """Max and mean |r2 - r2_fp64| over the sampled rows of one triangle:
    the engine's counts (exact integers) and r2 for those rows, against the
    FP64 closed form r2 = (c N - a b)^2 / (a (N - a) b (N - b)) evaluated in
    numpy on the same counts."""
    dres = eng.alloc(m * m * 4)
    dcnt = eng.alloc(m * m * 4)
    eng.correlate_dev(dA, dbc, m, dA, dbc, m, W, N, True, dres, m, dcnt, m)
    eng.sync()
    bc = dbc.download(m, np.uint32).astype(np.float64)
    errs = []
    for r in rows:
        j = np.arange(int(r) + 1, m)
        if j.size == 0:
            continue
        got = dres.download_at(int(r) * m * 4, m, np.float32)[j].astype(np.float64)
        c = dcnt.download_at(int(r) * m * 4, m, np.uint32)[j].astype(np.float64)
        a, b = bc[j], bc[int(r)]
        with np.errstate(divide="ignore", invalid="ignore"):
            want = (c * N - a * b) ** 2 / (a * (N - a) * b * (N - b))
        ok = np.isfinite(want) & (a > 0) & (b > 0)
        errs.append(np.abs(got[ok] - want[ok]))
    dres.free()
    dcnt.free()
    e = np.concatenate(errs)
    return {"max": float(e.max()), "mean": float(e.mean()), "cells": int(e.size),
            "rows": len(rows)}


def engine_grid(args):
    """Every contraction engine (popcount VALU, int8 MFMA, fp4 MFMA with
    direct operands) on single-region triangles of 64 to 30,000 SNPs at 256
    to 100,000 samples: kernel time (HIP events around the contraction
    kernel), Gpairs/s and the fraction of the engine's peak.  The evidence
    behind the AUTO engine rule."""
    import qld_amd
    rows = []
    for N in (256, 1000, 2048, 4096, 16622, 22554, 100000):
        for m in (64, 200, 500, 2000, 8000, 30000):
            A, bc = qld_amd.synth_blocks(20261203 + 60 + m, m, N)
            W = A.shape[1]
            pairs = m * (m - 1) // 2
            for engine in ("popcount", "mfma_i8", "mfma_fp4"):
                eng = qld_amd.Engine(engine=engine, r2_mode="exact", timing=True)
                dA = eng.alloc(A.nbytes).upload(A)
                dbc = eng.alloc(bc.nbytes).upload(bc)
                dres = eng.alloc(m * m * 4)
                for _ in range(10):     # clocks ramp over the first launches
                    eng.correlate_dev(dA, dbc, m, dA, dbc, m, W, N, True, dres, m)
                ks = []
                for _ in range(3):
                    eng.correlate_dev(dA, dbc, m, dA, dbc, m, W, N, True, dres, m)
                    ks.append(eng.last_contract_ms())
                for b in (dA, dbc, dres):
                    b.free()
                eng.close()
                k = float(np.median(ks))
                rl = roofline(engine, pairs, W, N, k / 1e3)
                rows.append({"samples": N, "snps": m, "engine": engine, "kernel_ms": round(k, 4),
                             "gpairs_s": round(pairs / k / 1e6, 2), "frac": round(rl["frac"], 3)})
                log(f"[bench] grid N={N} m={m} {engine}: {k:.4f} ms")
            best = min((r for r in rows if r["samples"] == N and r["snps"] == m),
                       key=lambda r: r["kernel_ms"])
            log(f"[bench] grid N={N} m={m}: fastest {best['engine']}")
    return rows


def config1_bench(args, steps):
    """Configuration 1: 2,000 SNPs x 16,622 samples, all-pairs r^2 (lower
    triangle) on the GPU -- the plain single-region call -- and on the CPU
    through the restatement of the BLIS path, all threads and one."""
    import oracle
    import qld_amd
    N, m = 16622, 2000
    A, bc = qld_amd.synth_blocks(20261203 + 1, m, N)
    W = A.shape[1]
    eng = qld_amd.Engine(engine="auto", r2_mode=args.r2_mode, timing=True)
    if args.ksplit is not None:
        eng.set_ksplit(args.ksplit)
    dA = eng.alloc(A.nbytes).upload(A)
    dbc = eng.alloc(bc.nbytes).upload(bc)
    dres = eng.alloc(m * m * 4)
    for _ in range(200):      # tiny launches: many to ramp the clocks
        eng.correlate_dev(dA, dbc, m, dA, dbc, m, W, N, True, dres, m)
    eng.sync()
    reps = max(steps, 100)
    eng.set_timing(False)     # timed calls without the per-call kernel events
    eng.timer_start()
    for _ in range(reps):
        eng.correlate_dev(dA, dbc, m, dA, dbc, m, W, N, True, dres, m)
    ms = eng.timer_stop() / reps
    eng.set_timing(True)
    kms = []
    for _ in range(10):
        eng.correlate_dev(dA, dbc, m, dA, dbc, m, W, N, True, dres, m)
        kms.append(eng.last_contract_ms())
    engine = eng.last_engine
    ksplit = eng.last_ksplit
    for b in (dA, dbc, dres):
        b.free()
    eng.close()
    pairs = m * (m - 1) // 2
    k_s = float(np.median(kms)) / 1e3
    cpu = host_cpu()
    threads = min(x for x in (cpu["physical_cores"], cpu["affinity_cpus"],
                              cpu["omp_num_threads"]) if x)
    out = {"workload": f"single region all-pairs r2, {m} SNPs x {N} samples, lower triangle",
           "pairs": pairs, "gpu": {"ms": round(ms, 4), "gpairs_s": round(pairs / ms / 1e6, 2),
                                   "engine": engine, "ksplit": ksplit,
                                   "roofline": roofline(
                                       engine, pairs, W, N, k_s,
                                       measured_traffic(f"single{m}_splitk", args.r2_mode)
                                       if ksplit == 3 else None, rows=m)}}
    for T in sorted({1, threads}):
        best = 1e30
        for _ in range(3):
            t0 = time.perf_counter()
            c = oracle.counts_blocked(A, tri=True, nthreads=T)
            oracle.r2_reference(bc, bc, c, N, tri=True)
            best = min(best, time.perf_counter() - t0)
        out[f"cpu_{T}_threads"] = {"s": round(best, 4), "gpairs_s": round(pairs / best / 1e9, 4)}
    out["gpu_vs_cpu_all_threads"] = round(out["gpu"]["gpairs_s"] /
                                          out[f"cpu_{threads}_threads"]["gpairs_s"], 1)
    log(f"[bench] config 1: GPU {ms:.4f} ms, CPU {out[f'cpu_{threads}_threads']['s']} s "
        f"on {threads} threads")
    return out


def union_pairs(starts, sizes, rows):
    """Pairs (i, j), i < j, that at least one window [s, s + size) holds:
    row i pairs with every j below the furthest end of a window starting at
    or before i (windows that end before i add nothing)."""
    best = np.full(rows, -1, np.int64)
    np.maximum.at(best, np.asarray(starts, np.int64),
                  np.asarray(starts, np.int64) + np.asarray(sizes, np.int64))
    best = np.maximum.accumulate(best)
    return int(np.maximum(best - np.arange(rows) - 1, 0).sum())


def windows_bench(args, steps):
    """4,096 windows, sizes round(exp(U(ln 50, ln 2000))), starts uniform,
    over a 500,000-SNP x 16,622-sample chromosome: one persistent launch
    (graph-replayed; the autotuner picks per-window or shared-pair work)
    against one launch per window."""
    import qld_amd
    N, rows, nwin = args.samples, args.win_rows, args.windows
    seed = 20261203 + 4
    log(f"[bench] windows: generating {rows} SNPs x {N} samples (seed {seed})")
    A, bc = qld_amd.synth_blocks(seed, rows, N)
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(50), np.log(2000), nwin))).astype(np.int64)
    starts = rng.integers(0, rows - sizes + 1)
    wins = qld_amd.window_layout(starts, sizes)
    total = int((sizes * sizes).sum())
    pairs = int((sizes * (sizes - 1) // 2).sum())
    eng = qld_amd.Engine(engine=args.engine, r2_mode=args.r2_mode, timing=True)
    t = eng.table(A, bc, N)
    out = eng.alloc(total * 4)
    del A
    if args.win_cutoff is None:
        log("[bench] windows: autotuning tile shapes")
        cutoff, _ = eng.autotune_windows(t, wins, out, reps=3)
    else:
        cutoff = args.win_cutoff
        eng.set_tile_cutoff(cutoff)
        eng.set_plan_schedule(args.win_sched)
        eng.set_plan_shared(args.win_shared)
    if args.no_graph:
        eng.set_plan_graphs(False)
    sched = eng.plan_schedule
    res = {"workload": f"{nwin} windows, sizes LogUniform(50, 2000), over {rows} SNPs x "
                       f"{N} samples, per-window lower-triangle r2",
           "windows": nwin, "pairs": pairs, "tile_cutoff": cutoff, "schedule": sched}

    def timed(plan, use_graph, k):
        # warm-up first: the GPU's clocks ramp over the first launches after
        # any idle gap (about ten launches of milliseconds each)
        plan.use_graph(use_graph)
        for _ in range(10):
            plan.run()
        eng.set_timing(False)     # timed runs without the per-call kernel events
        eng.sync()
        eng.timer_start()
        for _ in range(k):
            plan.run()
        ms = eng.timer_stop() / k
        eng.set_timing(True)
        return ms

    graphs = not args.no_graph
    shared_chosen = eng.plan_shared
    res["plan"] = "shared-pair" if shared_chosen else "per-window"

    def measure(mode):
        plan = eng.plan_windows(t, wins, out, mode)
        ms = timed(plan, graphs, steps)
        st = plan.stats()
        kms = []
        for _ in range(3):
            plan.run()
            kms.append(eng.last_contract_ms())
        plan.free()
        return ms, st, float(np.median(kms)) / 1e3

    # Per-window plan: every window's own wave blocks (the autotuned cutoff
    # and schedule).  Shared-pair plan: each pair some window holds counted
    # once over the chromosome's triangle, its r2 written into every window
    # holding it.
    eng.set_plan_shared(False)
    ms_w, st_w, k_w = measure(qld_amd.BATCH_PERSISTENT)
    ms_s, st_s, k_s_sh = measure(qld_amd.BATCH_SHARED)
    eng.set_plan_shared(shared_chosen)
    upairs = union_pairs(starts, sizes, rows)
    cover = np.zeros(rows + 1, np.int64)
    np.add.at(cover, starts, 1)
    np.add.at(cover, starts + sizes, -1)
    urows = int((np.cumsum(cover)[:rows] > 0).sum())     # rows some window holds
    traffic_ok = (nwin, rows, N) == (4096, 500000, 16622)
    per_window = {
        "ms": round(ms_w, 3), "gpairs_s": round(pairs / ms_w / 1e6, 2),
        "launches": st_w["launches"], "tiles": st_w["tiles"],
        "roofline": roofline("mfma_fp4", pairs, (N + 63) // 64, N, k_w,
                             measured_traffic(f"windows{nwin}", args.r2_mode)
                             if traffic_ok else None, rows=int(sizes.sum()))}
    shared = {
        "ms": round(ms_s, 3), "gpairs_s": round(pairs / ms_s / 1e6, 2),
        "units": st_s["tiles"], "launches": st_s["launches"], "distinct_pairs": upairs,
        "roofline": roofline("mfma_fp4", upairs, (N + 63) // 64, N, k_s_sh,
                             measured_traffic(f"windows{nwin}_shared", args.r2_mode)
                             if traffic_ok else None, rows=urows, out_pairs=pairs)}
    # The flop rate a per-window computation would need for the same outputs
    # in the shared plan's time: above the matrix cores' peak once pairs are
    # shared (the shared plan's own roofline counts distinct pairs only).
    shared["per_window_equivalent"] = {
        "achieved_tflop_s": round(pairs * 2.0 * N / k_s_sh / 1e12, 1),
        "frac_of_peak": round(pairs * 2.0 * N / k_s_sh / PEAK["mfma_fp4"], 3)}
    ms_b = ms_s if shared_chosen else ms_w
    st = st_s if shared_chosen else st_w
    # The baseline, one launch per window, on 128-tiles: a single window's
    # launch needs the finer tiles to fill the GPU (its fastest setting).
    ms_pg = ms_pd = None
    if not args.no_plain:
        eng.set_tile_cutoff(1 << 30)
        plain = eng.plan_windows(t, wins, out, qld_amd.BATCH_PLAIN)
        ms_pg = timed(plain, True, max(2, steps // 3)) if graphs else None
        ms_pd = timed(plain, False, max(2, steps // 3))
        plain.free()
    t.free()
    out.free()
    eng.close()
    res.update({"batched_ms": round(ms_b, 3), "batched_gpairs_s": round(pairs / ms_b / 1e6, 2),
                "batched_launches": st["launches"], "graphs": graphs,
                "roofline": (shared if shared_chosen else per_window)["roofline"],
                "per_window": per_window, "shared": shared})
    if ms_pd is not None:
        res.update({"plain_direct_ms": round(ms_pd, 3),
                    "plain_direct_gpairs_s": round(pairs / ms_pd / 1e6, 2),
                    "speedup_vs_plain": round(ms_pd / ms_b, 2)})
    if ms_pg is not None:
        res.update({"plain_graph_ms": round(ms_pg, 3),
                    "plain_graph_gpairs_s": round(pairs / ms_pg / 1e6, 2),
                    "speedup_vs_plain_graph": round(ms_pg / ms_b, 2)})
    log(f"[bench] windows: per-window plan {ms_w:.3f} ms, shared-pair plan {ms_s:.3f} ms "
        f"({upairs} distinct pairs); the autotuner chose {res['plan']}")
    log(f"[bench] windows: batched {ms_b:.3f} ms, plain graph {ms_pg} ms, "
        f"plain direct {ms_pd} ms (cutoff {cutoff}, schedule {sched}, graphs {graphs})")
    return res


def rect_bench(args, steps):
    """Region A 8,000 SNPs x region B 12,000 SNPs over 100,000 samples
    (Beta(0.5, 2) MAF, 5% of B copied from A with 2% bit flips): full
    rectangle of FP32 r2 tiles."""
    import qld_amd
    N, m, n = 100000, 8000, 12000
    seed = 20261203 + 3
    log(f"[bench] rect: generating {m} x {n} SNPs over {N} samples (seed {seed})")
    A, bA, B, bB = qld_amd.synth_regions(seed, m, n, N)
    W = A.shape[1]
    eng = qld_amd.Engine(engine=args.engine, r2_mode=args.r2_mode, timing=True)
    if args.tile:
        eng.set_plain_tile(args.tile)
    if args.fp4_operands:
        eng.set_fp4_operands(args.fp4_operands)
    dA = eng.alloc(A.nbytes).upload(A)
    dB = eng.alloc(B.nbytes).upload(B)
    dbA = eng.alloc(bA.nbytes).upload(bA)
    dbB = eng.alloc(bB.nbytes).upload(bB)
    dres = eng.alloc(m * n * 4)

    def step():
        eng.correlate_dev(dA, dbA, m, dB, dbB, n, W, N, False, dres, m)
    for _ in range(10):       # clocks ramp over the first launches
        step()
    eng.set_timing(False)     # timed steps without the per-call kernel events
    eng.sync()
    eng.timer_start()
    for _ in range(steps):
        step()
    ms = eng.timer_stop() / steps
    eng.set_timing(True)
    kms = []
    for _ in range(3):
        step()
        kms.append(eng.last_contract_ms())
    engine = eng.last_engine
    for b in (dA, dB, dbA, dbB, dres):
        b.free()
    eng.close()
    pairs = m * n
    k_s = float(np.mean(kms)) / 1e3
    log(f"[bench] rect: {ms:.3f} ms/step ({engine})")
    return {"workload": f"region A {m} SNPs x region B {n} SNPs over {N} samples, "
                        "full rectangle FP32 r2", "pairs": pairs, "engine": engine,
            "ms_per_step": round(ms, 3), "gpairs_s": round(pairs / ms / 1e6, 2),
            "roofline": roofline(engine, pairs, W, N, k_s,
                                 measured_traffic("rect100k", args.r2_mode), rows=m + n)}


def report_bench(args, A, bc, N):
    """End to end through the drop-in report call: host tables in, report text
    out (to /dev/null), r2limit 0.2, single region.  Device-side selection with
    streamed hits (qld_correlate_report) vs the full r^2 matrix copied to the
    host and scanned there (qld_gpu_correlate + qld_write_results, the
    reference's correlate_gpu -> writeResults flow)."""
    import ctypes as C
    import qld_amd
    m = A.shape[0]
    W = A.shape[1]
    pos = (np.arange(m, dtype=np.uint32) * 50 + 10000)
    ids = [f"rs{i}" for i in range(m)]
    idarr = (C.c_char_p * m)(*[s.encode() for s in ids])
    idp = C.cast(idarr, C.c_void_p)
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    eng = qld_amd.Engine(engine="auto", r2_mode="reference")
    host = qld_amd.host_lib()
    lim = 0.2

    def streamed():
        fp = libc.fopen(b"/dev/null", b"w")
        t0 = time.perf_counter()
        rc = host.qld_correlate_report(eng._ctx, fp, A.ctypes.data, pos.ctypes.data, idp,
                                       bc.ctypes.data, m, A.ctypes.data, pos.ctypes.data, idp,
                                       bc.ctypes.data, m, W, N, 0, C.c_float(lim))
        libc.fclose(fp)
        assert rc == 0
        return time.perf_counter() - t0

    def full_matrix():
        fp = libc.fopen(b"/dev/null", b"w")
        t0 = time.perf_counter()
        res, _ = eng.correlate(A, bc, snp_size=N, triangular=True)
        host.qld_write_results_header(fp)
        lines = host.qld_write_results(fp, pos.ctypes.data, idp, bc.ctypes.data, m,
                                       pos.ctypes.data, idp, bc.ctypes.data, m, N,
                                       res.ctypes.data, C.c_float(lim), 0)
        libc.fclose(fp)
        return time.perf_counter() - t0, lines

    cli = cli_bench(A, bc, N, pos, ids)

    log("[bench] report: streamed device selection")
    streamed()
    ts = min(streamed() for _ in range(2))
    log("[bench] report: full matrix to host")
    tf, lines = full_matrix()
    eng.close()
    pairs = m * (m - 1) // 2
    log(f"[bench] report {m} SNPs: streamed {ts:.3f} s, full matrix {tf:.3f} s, {lines} lines")
    return {"workload": f"drop-in report, {m} SNPs x {N} samples, r2limit {lim}, text to "
                        "/dev/null, host tables in",
            "lines": int(lines), "pairs": pairs,
            "streamed_s": round(ts, 3), "full_matrix_s": round(tf, 3),
            "streamed_gpairs_s": round(pairs / ts / 1e9, 3),
            "speedup": round(tf / ts, 2), "cli": cli}


def cli_bench(A, bc, N, pos, ids, chunk=5000):
    """qLD-compute on an MDF directory of the same table (chunks of 5,000
    SNPs, written by qld_mdf_write): wall time of the whole command, report
    to /dev/null, and the MDF ingest alone (-dump-tables)."""
    import gzip
    import shutil
    import subprocess
    import tempfile
    import qld_amd
    m = A.shape[0]
    d = tempfile.mkdtemp(prefix="qld_mdf_")
    try:
        mdf = os.path.join(d, "chr1")
        os.makedirs(mdf)
        hdr = ("##fileformat=VCFv4.2\n#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t"
               + "\t".join(f"S{i}" for i in range(N)) + "\n"
               + f"chr1 {chunk} {N} {m} {int(pos[0])} {int(pos[-1])}\n")
        with gzip.open(os.path.join(mdf, "chr1_header.vcf.gz"), "wb") as f:
            f.write(hdr.encode())
        for a in range(0, m, chunk):
            b = min(a + chunk, m)
            qld_amd.mdf_write(os.path.join(mdf, f"chr1_{a + 1}_{b}_{int(pos[a])}_{int(pos[b - 1])}.mdf"),
                              "chr1", A[a:b], pos[a:b], ids[a:b], bc[a:b])
        mb = sum(os.path.getsize(os.path.join(mdf, f)) for f in os.listdir(mdf)) / 1e6
        exe = os.path.join(ROOT, "qld_amd", "lib", "qLD-compute")
        base = [exe, "-input", mdf, "-ploidy", "haploid", "-mdf", "-gpu", "-r2limit", "0.2"]
        t0 = time.perf_counter()
        r = subprocess.run(base + ["-output", os.path.join(d, "t.txt"), "-dump-tables"],
                           cwd=d, capture_output=True, text=True, timeout=600)
        t_ingest = time.perf_counter() - t0
        assert r.returncode == 0, r.stderr
        t0 = time.perf_counter()
        r = subprocess.run(base + ["-output", "/dev/null"], cwd=d, capture_output=True,
                           text=True, timeout=600)
        t_all = time.perf_counter() - t0
        assert r.returncode == 0, r.stderr
    finally:
        shutil.rmtree(d, ignore_errors=True)
    log(f"[bench] qLD-compute {m} SNPs: {t_all:.3f} s end to end, ingest {t_ingest:.3f} s "
        f"({mb:.0f} MB MDF)")
    return {"mdf_mb": round(mb, 1), "wall_s": round(t_all, 3),
            "ingest_and_dump_s": round(t_ingest, 3)}


def list_bench(args):
    """qLD-compute -inputList: 1,024 single-region windows (sizes LogUniform(50,
    2000) SNPs) over one 100,000-SNP x 16,622-sample MDF chromosome, batched
    into one persistent plan vs run task by task (-plain-tasks); every report
    to its own file at r2limit 0.2."""
    import gzip
    import shutil
    import subprocess
    import tempfile
    import qld_amd
    N, m, nwin = 16622, args.list_snps, args.list_windows
    A, bc = qld_amd.synth_blocks(args.seed + 7, m, N)
    pos = (np.arange(m, dtype=np.uint32) * 50 + 10000)
    ids = [f"rs{i}" for i in range(m)]
    rng = np.random.default_rng(args.seed + 8)
    sizes = np.exp(rng.uniform(np.log(50), np.log(2000), nwin)).round().astype(np.int64)
    starts = rng.integers(0, m - sizes + 1)
    pairs = int((sizes * (sizes - 1) // 2).sum())
    d = tempfile.mkdtemp(prefix="qld_list_")
    try:
        mdf = os.path.join(d, "chr1")
        os.makedirs(mdf)
        chunk = 5000
        hdr = ("##fileformat=VCFv4.2\n#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t"
               + "\t".join(f"S{i}" for i in range(N)) + "\n"
               + f"chr1 {chunk} {N} {m} {int(pos[0])} {int(pos[-1])}\n")
        with gzip.open(os.path.join(mdf, "chr1_header.vcf.gz"), "wb") as f:
            f.write(hdr.encode())
        for a in range(0, m, chunk):
            b = min(a + chunk, m)
            qld_amd.mdf_write(os.path.join(mdf, f"chr1_{a + 1}_{b}_{int(pos[a])}_{int(pos[b - 1])}.mdf"),
                              "chr1", A[a:b], pos[a:b], ids[a:b], bc[a:b])
        with open(os.path.join(d, "list.csv"), "w") as f:
            for s0, sz in zip(starts, sizes):
                lo, hi = int(pos[s0]), int(pos[s0 + sz - 1])
                f.write(f"{mdf},{lo},{hi},{mdf},{lo},{hi}\n")
        exe = os.path.join(ROOT, "qld_amd", "lib", "qLD-compute")
        out = {}
        for name, extra in (("batched", []), ("plain_tasks", ["-plain-tasks"])):
            od = os.path.join(d, name)
            os.makedirs(od)
            log(f"[bench] inputList {nwin} windows: {name}")
            t0 = time.perf_counter()
            r = subprocess.run([exe, "-inputList", os.path.join(d, "list.csv"), "-ploidy",
                                "haploid", "-mdf", "-gpu", "-r2limit", "0.2", "-output",
                                os.path.join(od, "w.txt")] + extra, cwd=d,
                               capture_output=True, text=True, timeout=900)
            out[name] = round(time.perf_counter() - t0, 3)
            assert r.returncode == 0, r.stderr[-2000:]
        same = all(open(os.path.join(d, "batched", f), "rb").read()
                   == open(os.path.join(d, "plain_tasks", f), "rb").read()
                   for f in os.listdir(os.path.join(d, "batched")))
    finally:
        shutil.rmtree(d, ignore_errors=True)
    log(f"[bench] inputList: batched {out['batched']} s, task by task {out['plain_tasks']} s, "
        f"reports identical: {same}")
    return {"workload": f"qLD-compute -inputList, {nwin} windows LogUniform(50, 2000) SNPs over "
                        f"{m} SNPs x {N} samples (MDF, chunks of 5,000), r2limit 0.2, one report "
                        "file per window",
            "pairs": pairs, "batched_s": out["batched"], "plain_tasks_s": out["plain_tasks"],
            "speedup": round(out["plain_tasks"] / out["batched"], 2),
            "reports_identical": bool(same)}


def serving_bench(args):
    """Poisson arrivals of LD queries (window size LogUniform(20, 5000), a
    uniformly chosen resident chromosome, uniform start) against three
    16,622-sample and two 200,000-sample chromosomes; per rate: p50/p99
    latency (submit -> r2 ready in HBM) and pairs/s."""
    import qld_amd
    rates = [float(x) for x in args.serving_rates.split(",")]
    nreq = args.serving_requests
    rows16, rows200 = args.serving_rows16, args.serving_rows200
    eng = qld_amd.Engine(engine="auto", r2_mode=args.r2_mode)
    tabs = []
    specs = [(16622, rows16)] * 3 + [(200000, rows200)] * 2
    for k, (N, rows) in enumerate(specs):
        log(f"[bench] serving: chromosome {k}: {rows} SNPs x {N} samples")
        A, bc = qld_amd.synth_blocks(20261203 + 50 + k, rows, N)
        tabs.append((eng.table(A, bc, N), rows, N))
        del A
    srv = qld_amd.Server(eng, out_budget=8 << 30, max_batch=128, streams=3)
    rng = np.random.default_rng(20261203 + 5)
    results = []
    # warm-up
    ids = [srv.submit(tabs[0][0], 0, 1000, qld_amd.REQ_DISCARD) for _ in range(4)]
    for i in ids:
        srv.wait(i, 120000.0)
        srv.release(i)
    for rate in rates:
        gaps = rng.exponential(1.0 / rate, nreq)
        which = rng.integers(0, len(tabs), nreq)
        sizes = np.rint(np.exp(rng.uniform(np.log(20), np.log(5000), nreq))).astype(int)
        sizes = np.minimum(sizes, [tabs[w][1] for w in which])
        starts = [int(rng.integers(0, tabs[w][1] - z + 1)) for w, z in zip(which, sizes)]
        ids = []
        b0 = srv.stats()["batches"]
        t0 = time.perf_counter()
        due = t0
        for k in range(nreq):
            due += gaps[k]
            while time.perf_counter() < due:
                pass
            ids.append(srv.submit(tabs[which[k]][0], starts[k], int(sizes[k]),
                                  qld_amd.REQ_DISCARD))
        lat = []
        for i in ids:
            lat.append(srv.wait(i, 120000.0))
            srv.release(i)
        span = time.perf_counter() - t0
        nb = srv.stats()["batches"] - b0
        lat = np.array(lat)
        pairs = int((sizes * (sizes - 1) // 2).sum())
        pairs16 = int(sum(z * (z - 1) // 2 for w, z in zip(which, sizes) if tabs[w][2] == 16622))
        res = {"rate_req_s": rate, "requests": nreq,
               "achieved_req_s": round(nreq / span, 1),
               "p50_ms": round(float(np.percentile(lat, 50)), 3),
               "p99_ms": round(float(np.percentile(lat, 99)), 3),
               "mean_ms": round(float(lat.mean()), 3),
               "gpairs_s": round(pairs / span / 1e9, 3),
               "gpairs_s_16622": round(pairs16 / span / 1e9, 3),
               "batches": nb, "requests_per_batch": round(nreq / max(nb, 1), 2)}
        log(f"[bench] serving {rate:g} req/s: p50 {res['p50_ms']} ms, p99 {res['p99_ms']} ms, "
            f"{res['achieved_req_s']} req/s, {res['requests_per_batch']} requests per batch")
        results.append(res)
    st = srv.stats()
    srv.close()
    for t, _, _ in tabs:
        t.free()
    eng.close()
    return {"workload": f"Poisson LD queries, windows LogUniform(20, 5000) over 3 x {rows16} "
                        f"SNPs x 16,622 and 2 x {rows200} SNPs x 200,000 samples, resident; "
                        "latency = submit to r2 ready in HBM",
            "rates": results, "batches": st["batches"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 50 timed steps of the single region (0.11 s): the pipeline's start-up
    # (about 0.3 ms) adds 6 us per step to the wall-clock figure, 15 us at 20.
    ap.add_argument("--steps", type=int, default=50)
    # The first launches after idle run up to 25% slower while the GPU's
    # clocks ramp up (2.67 vs 2.61 ms/step at 2 vs 10 warm-up steps).
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--snps", type=int, default=30000)
    ap.add_argument("--samples", type=int, default=16622)
    ap.add_argument("--engine", default="auto")
    ap.add_argument("--r2-mode", default="exact",
                    help="exact: FP32 r2 within one ulp of the FP64 closed form (default); "
                         "reference: bit-identical to the reference's float/double formula")
    ap.add_argument("--seed", type=int, default=20261203 + 2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-windows", action="store_true")
    ap.add_argument("--no-rect", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-report", action="store_true")
    ap.add_argument("--fp4-operands", default="",
                    help="fp4 operand source: expanded | packed | direct (engine default if empty)")
    ap.add_argument("--tile", type=int, default=0,
                    help="tile edge of the single-region MFMA launch (0: engine default)")
    ap.add_argument("--no-list", action="store_true")
    ap.add_argument("--ksplit", type=int, default=None,
                    help="split-K chunks for the configuration-1 leg (0 by shape, 1 none)")
    ap.add_argument("--list-snps", type=int, default=100000)
    ap.add_argument("--list-windows", type=int, default=1024)
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run one device copy of this many MiB (a known-size "
                         "kernel for calibrating HBM traffic counters)")
    # the configuration's sweep (100 to 5,000 requests/s) and one rate past it,
    # where requests queue and batches grow
    ap.add_argument("--serving-rates", default="100,200,500,1000,2000,5000,20000")
    ap.add_argument("--serving-requests", type=int, default=2000)
    ap.add_argument("--serving-rows16", type=int, default=100000)
    ap.add_argument("--serving-rows200", type=int, default=20000)
    ap.add_argument("--windows", type=int, default=4096)
    ap.add_argument("--win-rows", type=int, default=500000)
    ap.add_argument("--win-cutoff", type=int, default=None,
                    help="window plans: tile cutoff to use instead of autotuning")
    ap.add_argument("--win-sched", default="waves", choices=["waves", "tiles"],
                    help="window plans: work schedule with --win-cutoff")
    ap.add_argument("--win-shared", action="store_true",
                    help="with --win-cutoff: the batched plan is the shared-pair form")
    ap.add_argument("--no-graph", action="store_true",
                    help="window plans: direct launches, no HIP graph capture")
    ap.add_argument("--no-plain", action="store_true",
                    help="window plans: skip the one-launch-per-window baseline")
    ap.add_argument("--engine-grid", action="store_true",
                    help="only time every engine on a grid of shapes and print that table")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(each box has one MI355X; the path would shard by output tiles)",
              file=sys.stderr)
        sys.exit(2)

    import qld_amd

    if args.engine_grid:
        print(json.dumps({"engine_grid": engine_grid(args)}), flush=True)
        return

    N, m = args.samples, args.snps
    log(f"[bench] generating {m} SNPs x {N} samples (seed {args.seed})")
    A, bc = qld_amd.synth_blocks(args.seed, m, N)
    W = A.shape[1]
    pairs = m * (m - 1) // 2

    eng = qld_amd.Engine(engine=args.engine, r2_mode=args.r2_mode, timing=True)
    if args.tile:
        eng.set_plain_tile(args.tile)
    if args.fp4_operands:
        eng.set_fp4_operands(args.fp4_operands)
    dA = eng.alloc(A.nbytes).upload(A)
    dbc = eng.alloc(bc.nbytes).upload(bc)
    dres = eng.alloc(m * m * 4)

    def step():
        eng.correlate_dev(dA, dbc, m, dA, dbc, m, W, N, True, dres, m)

    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    eng.sync()
    engine = eng.last_engine

    log(f"[bench] timing {args.steps} steps, engine {engine}")
    # The timed steps run as a caller would run them, without the per-call
    # kernel events; the kernel is timed afterwards with them.
    eng.set_timing(False)
    eng.sync()
    t0 = time.perf_counter()
    eng.timer_start()
    for _ in range(args.steps):
        step()
    dev_ms = eng.timer_stop()
    eng.sync()
    wall_s = time.perf_counter() - t0
    ms_per_step = wall_s * 1e3 / args.steps
    eng.set_timing(True)

    # Per-launch duration of the contraction kernel (events around it alone).
    kms = []
    for _ in range(max(3, min(args.steps, 10))):
        step()
        kms.append(eng.last_contract_ms())
    kernel_s = float(np.mean(kms)) / 1e3
    log(f"[bench] {ms_per_step:.3f} ms/step wall, {dev_ms / args.steps:.3f} ms/step "
        f"device, kernel {kernel_s * 1e3:.3f} ms")

    value = pairs / (ms_per_step / 1e3) / 1e9

    # The other r2 mode on the same data and engine, for comparison.
    other = "reference" if args.r2_mode == "exact" else "exact"
    eng2 = qld_amd.Engine(engine=args.engine, r2_mode=other, timing=True)
    if args.tile:
        eng2.set_plain_tile(args.tile)
    if args.fp4_operands:
        eng2.set_fp4_operands(args.fp4_operands)
    dA2 = eng2.alloc(A.nbytes).upload(A)
    dbc2 = eng2.alloc(bc.nbytes).upload(bc)
    dres2 = eng2.alloc(m * m * 4)
    for _ in range(args.warmup):
        eng2.correlate_dev(dA2, dbc2, m, dA2, dbc2, m, W, N, True, dres2, m)
    eng2.set_timing(False)
    eng2.sync()
    eng2.timer_start()
    for _ in range(args.steps):
        eng2.correlate_dev(dA2, dbc2, m, dA2, dbc2, m, W, N, True, dres2, m)
    other_ms = eng2.timer_stop() / args.steps
    eng2.set_timing(True)
    sample_rows = np.sort(np.random.default_rng(64).choice(m, min(64, m), replace=False))
    err_other = r2_error_vs_fp64(eng2, dA2, dbc2, m, W, N, sample_rows)
    for b in (dA2, dbc2, dres2):
        b.free()
    eng2.close()
    log(f"[bench] r2 mode {other}: {other_ms:.3f} ms/step device")
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gpairs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": DTYPE.get(engine, engine),
        "data": f"synthetic haplotype-block Markov table, seed {args.seed}",
        "config": {"workload": f"single region all-pairs r2, {m} SNPs x {N} haploid "
                               "samples, lower triangle, FP32 r2 tiles in HBM",
                   "snps": m, "samples": N, "words": W, "pairs": pairs,
                   "engine": engine, "r2_mode": args.r2_mode,
                   "parallelism": "1 GPU"},
        "device_ms_per_step": round(dev_ms / args.steps, 4),
        f"{other}_mode": {"ms_per_step": round(other_ms, 4),
                          "gpairs_s": round(pairs / other_ms / 1e6, 3)},
        "r2_error_vs_fp64": {args.r2_mode: r2_error_vs_fp64(eng, dA, dbc, m, W, N, sample_rows),
                             other: err_other},
        "roofline": roofline(engine, pairs, W, N, kernel_s,
                             measured_traffic(f"single{m // 1000}k", args.r2_mode)
                             if (m, N) == (30000, 16622) and not args.fp4_operands else None,
                             rows=m),
    }
    if args.copy_probe:
        nb = args.copy_probe << 20
        src, dst = eng.alloc(nb).fill(7), eng.alloc(nb)
        eng._check(qld_amd.hip_lib().qld_dev_copy(eng._ctx, dst.ptr, src.ptr, nb))
        eng.sync()
        src.free()
        dst.free()
    for b in (dA, dbc, dres):
        b.free()
    eng.close()

    if not args.no_windows:
        out["batched_windows"] = windows_bench(args, max(3, min(args.steps, 20)))
    if not args.no_rect:
        out["rect_100k"] = rect_bench(args, max(3, min(args.steps, 20) // 2))
    if not args.no_serving:
        out["serving"] = serving_bench(args)
    if not args.no_report:
        out["report_end_to_end"] = report_bench(args, A, bc, N)
    if not args.no_list:
        out["input_list"] = list_bench(args)

    if not args.no_cpu:
        log("[bench] config 1 (2,000 SNPs): GPU and CPU")
        out["config1"] = config1_bench(args, max(3, min(args.steps, 20)))
        log("[bench] CPU baseline (oracle restatement of the BLIS path)")
        out["cpu_baseline"] = cpu_baseline(A, bc, N)
        out["speedup_vs_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
