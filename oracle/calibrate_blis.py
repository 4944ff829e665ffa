"""Calibrate the CPU baseline against the reference's own BLIS CPU path.

TEST INFRASTRUCTURE, run only in the build container (it needs the
reference checkout and `make -C oracle ref-blis ref`):

1. Parity of the reference's BLIS build: every single-task golden CLI case
   (reports written by the reference's non-BLIS build, tests/golden/cli/)
   is rerun with `-blis` and must be byte-identical -- integer exactness of
   the popcount counts through BLIS's double-typed syrk/gemm edge paths
   (bli_gemm_ker_var2.c:289-317, bli_herk_l_ker_var2.c:340-377).
2. Timing on the benchmark's single-region data (configuration 1: 2,000 SNPs,
   configuration 2: 30,000 SNPs, 16,622 samples, the benchmark's generator
   and seeds) written as an MDF directory: the reference's `-blis` run with
   BLIS threads jc x ic = 8 (bli_gemm_threading.c:116-121) and with 1, its
   own "[T] GEMM times" + "[T] LD times" (gemm_cpu.c:593-598: the counts and
   the r^2 epilogue, the region the oracle restatement times), against the
   oracle's blocked restatement (qo_counts_blocked + qo_r2_reference) on the
   same rows with the same thread counts.

    python oracle/calibrate_blis.py [--snps 2000,30000] [--threads 8] [--write]

Prints one JSON object per line; --write also stores the timings in
oracle/blis_calibration.json, which bench.py's CPU baseline reports.
"""
from __future__ import annotations

import argparse
import glob
import gzip
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
import qld_amd  # noqa: E402

BLIS_BIN = os.path.join(HERE, "_ref", "qLD-compute-blis")
CLI = os.path.join(ROOT, "tests", "golden", "cli")


def parity():
    cases = json.load(open(os.path.join(CLI, "cases.json")))["cases"]
    out = {}
    for name, args in cases.items():
        if name == "input_list":
            continue
        with tempfile.TemporaryDirectory() as work:
            os.symlink(os.path.join(CLI, "data"), os.path.join(work, "data"))
            os.makedirs(os.path.join(work, "out"))
            r = subprocess.run([BLIS_BIN] + args + ["-blis", "-mdf", "-threads", "1",
                                                    "-output", "out/report.txt"],
                               cwd=work, capture_output=True, text=True, timeout=300,
                               env=dict(os.environ, BLIS_JC_NT="1", BLIS_IC_NT="1"))
            got = open(os.path.join(work, "out", "report.txt"), "rb").read() \
                if r.returncode == 0 else None
        want = open(os.path.join(CLI, "expected", name, "report.txt"), "rb").read()
        out[name] = got == want
    return out


def write_mdf(d, A, bc, N, chunk=5000):
    m = A.shape[0]
    pos = (np.arange(m, dtype=np.uint32) * 50 + 10000)
    ids = [f"rs{i}" for i in range(m)]
    os.makedirs(d)
    hdr = ("##fileformat=VCFv4.2\n#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t"
           + "\t".join(f"S{i}" for i in range(N)) + "\n"
           + f"chr1 {chunk} {N} {m} {int(pos[0])} {int(pos[-1])}\n")
    with gzip.open(os.path.join(d, "chr1_header.vcf.gz"), "wb") as f:
        f.write(hdr.encode())
    for a in range(0, m, chunk):
        b = min(a + chunk, m)
        qld_amd.mdf_write(os.path.join(d, f"chr1_{a + 1}_{b}_{int(pos[a])}_{int(pos[b - 1])}.mdf"),
                          "chr1", A[a:b], pos[a:b], ids[a:b], bc[a:b])


def reference_time(d, jc, ic):
    """The reference's own GEMM + LD phase times of one -blis run."""
    with tempfile.TemporaryDirectory() as work:
        env = dict(os.environ, BLIS_JC_NT=str(jc), BLIS_IC_NT=str(ic),
                   OMP_NUM_THREADS=str(jc * ic))
        t0 = time.perf_counter()
        r = subprocess.run([BLIS_BIN, "-input", d, "-mdf", "-blis", "-ploidy", "haploid",
                            "-threads", "1", "-r2limit", "2", "-output", "out.txt"],
                           cwd=work, capture_output=True, text=True, timeout=3600, env=env)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise SystemExit(r.stderr[-2000:])
        text = "".join(open(p).read() for p in glob.glob(os.path.join(work, "*"))
                       if os.path.isfile(p) and not p.endswith("out.txt"))
    gemm = sum(float(x) for x in re.findall(r"\[T\] GEMM times: ([0-9.eE+-]+)", text))
    ld = sum(float(x) for x in re.findall(r"\[T\] LD times: ([0-9.eE+-]+)", text))
    return gemm, ld, wall


def restatement_time(A, bc, N, threads, reps=3):
    best = 1e30
    for _ in range(reps):
        t0 = time.perf_counter()
        c = oracle.counts_blocked(A, tri=True, nthreads=threads)
        oracle.r2_reference(bc, bc, c, N, tri=True)
        best = min(best, time.perf_counter() - t0)
        del c
    return best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--snps", default="2000,30000")
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--write", action="store_true")
    args = ap.parse_args()
    record = {"host": cpu_model(), "restatement_isa_bits": oracle.counts_blocked_isa(),
              "configs": []}
    if not os.path.exists(BLIS_BIN):
        raise SystemExit("build it first: make -C oracle ref-blis")
    if not args.no_parity:
        p = parity()
        record["blis_reports_identical"] = p
        print(json.dumps({"blis_reports_identical": p, "all": all(p.values())}), flush=True)
    N = 16622
    for m in (int(x) for x in args.snps.split(",")):
        seed = 20261203 + (1 if m == 2000 else 2)
        A, bc = qld_amd.synth_blocks(seed, m, N)
        with tempfile.TemporaryDirectory() as tmp:
            d = os.path.join(tmp, "chr1")
            write_mdf(d, A, bc, N)
            res = {"snps": m, "samples": N, "pairs": m * (m - 1) // 2}
            for T in sorted({1, args.threads}):
                jc = 2 if T % 2 == 0 and T > 1 else 1
                gemm, ld, wall = reference_time(d, jc, T // jc)
                mine = restatement_time(A, bc, N, T)
                res[f"threads_{T}"] = {
                    "reference_blis_gemm_s": round(gemm, 3), "reference_ld_s": round(ld, 3),
                    "reference_counts_r2_s": round(gemm + ld, 3),
                    "reference_wall_s": round(wall, 3),
                    "restatement_s": round(mine, 3),
                    "restatement_speedup": round((gemm + ld) / mine, 2)}
                print(json.dumps(res), flush=True)
            record["configs"].append(res)
    if args.write:
        with open(os.path.join(HERE, "blis_calibration.json"), "w") as f:
            json.dump(record, f, indent=1)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip() + f", {os.cpu_count()} cpus"
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
