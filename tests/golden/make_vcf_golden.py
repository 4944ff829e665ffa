"""Write the qLD-compute VCF-input fixtures under tests/golden/vcf/.

1. Seeded VCF chunk directories (``vcf/data/<id>/``) in the parser's layout
   (parse_vcf_output.c:16-200): ``<id>_header.vcf.gz`` holds the VCF header
   lines, the ``#CHROM`` line and then ``id snipsPerFile snipSize totalSnips
   posMin posMax`` (parse_vcf.c:620-634); each chunk
   ``<id>_<first>_<last>_<posMin>_<posMax>.vcf.gz`` repeats the first header
   line and the ``#CHROM`` line before its data lines.  The data lines mix
   every kind of row the correlator filters (multi-base REF/ALT, FILTER not
   PASS, AF 0 or 1, VT not SNP, FORMAT without GT, monomorphic sites, N
   before / after the second base, ALT ".") with kept biallelic SNPs; GT is
   haploid, phased diploid or unphased diploid (whose alleles the reference
   shuffles with rand()).
2. For each case, the report and ``input_report.txt`` written by the
   reference correlator's own CPU build (``make -C oracle ref-cli``), run
   here without -mdf -> ``vcf/expected/<case>/``.

Only this container has the reference checkout; everything written is data.
Usage: python tests/golden/make_vcf_golden.py
"""
from __future__ import annotations

import gzip
import json
import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "vcf")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "qLD-compute")
BASES = "ACGT"
LINE1 = "##fileformat=VCFv4.2"


def gt_text(alleles, ploidy, phased):
    if ploidy == 1:
        return alleles[0]
    return ("|" if phased else "/").join(alleles)


def make_rows(rng, nrows, nsamp, ploidy, phased, start_pos):
    """Data lines (without CHROM, filled in later) with LD structure and a
    share of every filtered kind."""
    rows = []
    pos = start_pos
    prev = rng.integers(0, 2, size=(nsamp, ploidy))
    for r in range(nrows):
        pos += int(rng.integers(1, 30))
        ref, alt = rng.choice(list(BASES), 2, replace=False)
        kind = rng.random()
        filt, info, fmt = "PASS", f"AF={rng.uniform(0.05, 0.95):.3f};VT=SNP", "GT"
        # LD: copy the previous site's alleles with probability 0.8
        fresh = rng.integers(0, 2, size=(nsamp, ploidy))
        keep = rng.random((nsamp, ploidy)) < 0.8
        g = np.where(keep, prev, fresh)
        prev = g
        calls = [[str(x) for x in g[s]] for s in range(nsamp)]
        extra = ""
        if kind < 0.05:
            ref = ref + rng.choice(list(BASES))             # multi-base REF
        elif kind < 0.09:
            alt = alt + "," + rng.choice([b for b in BASES if b not in (ref, alt)])
        elif kind < 0.13:
            filt = "LowQual"
        elif kind < 0.16:
            info = "AF=0;VT=SNP"
        elif kind < 0.19:
            info = "AF=1"
        elif kind < 0.22:
            info = "AF=0.2;VT=INDEL"
        elif kind < 0.26:
            info = "."                                      # AF taken as 0.9
        elif kind < 0.29:
            info = "DP=20"                                  # no AF, no VT: kept
        elif kind < 0.32:
            info = "AF=0,0.37;VT=SNP"                       # first AF in (0,1)
        elif kind < 0.35:
            fmt = "DP"                                      # no GT: dropped
        elif kind < 0.40:
            fmt = "GT:DP"
            extra = ":7"
        elif kind < 0.44:
            calls = [["0"] * ploidy for _ in range(nsamp)]  # monomorphic: dropped
        elif kind < 0.47:
            calls[0][0] = "."                               # N first: dropped
        elif kind < 0.50:
            calls[-1][-1] = "."                             # N after both bases: kept
        elif kind < 0.53:
            alt = "."                                       # ALT "." -> REF
        elif kind < 0.55:
            filt = "."
        samples = [gt_text(c, ploidy, phased) + extra for c in calls]
        rows.append((pos, f"snp{r}", ref, alt, "50", filt, info, fmt, samples))
    return rows


def write_dir(name, rows, nsamp, ploidy, spf):
    d = os.path.join(OUT, "data", name)
    os.makedirs(d, exist_ok=True)
    chrom = "#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t" + "\t".join(
        f"{name}S{s}" for s in range(nsamp))
    n = len(rows)
    snip = nsamp * ploidy
    hdr = (f"{LINE1}\n##source=qld_amd VCF fixture\n##INFO=<ID=AF,Number=A,Type=Float>\n"
           f"{chrom}\n{name} {spf} {snip} {n} {rows[0][0]} {rows[-1][0]}\n")
    with gzip.GzipFile(os.path.join(d, f"{name}_header.vcf.gz"), "wb", mtime=0) as f:
        f.write(hdr.encode())
    for a in range(0, n, spf):
        b = min(a + spf, n)
        lines = [LINE1, chrom]
        for (pos, rid, ref, alt, q, filt, info, fmt, samples) in rows[a:b]:
            lines.append("\t".join([name, str(pos), rid, ref, alt, q, filt, info, fmt] + samples))
        fn = f"{name}_{a + 1}_{b}_{rows[a][0]}_{rows[b - 1][0]}.vcf.gz"
        with gzip.GzipFile(os.path.join(d, fn), "wb", mtime=0) as f:
            f.write(("\n".join(lines) + "\n").encode())
    return rows


def build():
    rng = np.random.default_rng(20261204)
    shutil.rmtree(os.path.join(OUT, "data"), ignore_errors=True)
    P = {}
    P["vh"] = write_dir("vh", make_rows(rng, 90, 45, 1, True, 1000), 45, 1, 30)
    P["vh2"] = write_dir("vh2", make_rows(rng, 60, 45, 1, True, 5000), 45, 1, 25)
    P["vp"] = write_dir("vp", make_rows(rng, 70, 30, 2, True, 2000), 30, 2, 35)
    P["vu"] = write_dir("vu", make_rows(rng, 80, 30, 2, False, 3000), 30, 2, 20)
    P["vu2"] = write_dir("vu2", make_rows(rng, 50, 30, 2, False, 9000), 30, 2, 50)
    with open(os.path.join(OUT, "data", "samples_vu.txt"), "w") as f:
        f.write("\n".join(f"vuS{s}" for s in range(0, 30, 3)) + "\n")
    with open(os.path.join(OUT, "data", "samples_vh.txt"), "w") as f:
        f.write(" ".join(f"vhS{s}" for s in range(5, 40)) + "\n")
    return {k: [r[0] for r in v] for k, v in P.items()}


def cases(P):
    h, u = P["vh"], P["vu"]
    return {
        "haploid_all": ["-input", "data/vh", "-ploidy", "haploid", "-r2limit", "0.05"],
        "haploid_window": ["-input", "data/vh", "-ploidy", "haploid", "-posWmin1", str(h[10]),
                           "-posWmax1", str(h[70]), "-r2limit", "0.1"],
        "haploid_samples": ["-input", "data/vh", "-ploidy", "haploid", "-sampleList",
                            "data/samples_vh.txt", "-r2limit", "0.05"],
        "haploid_two_inputs": ["-input", "data/vh", "-input2", "data/vh2", "-ploidy", "haploid",
                               "-posWmin1", str(h[0]), "-posWmax1", str(h[60]),
                               "-posWmin2", str(P["vh2"][2]), "-posWmax2", str(P["vh2"][50]),
                               "-r2limit", "0.02"],
        "phased": ["-input", "data/vp", "-ploidy", "phased_diploid", "-r2limit", "0.05"],
        "unphased": ["-input", "data/vu", "-ploidy", "unphased_diploid", "-r2limit", "0.05"],
        "unphased_two_windows": ["-input", "data/vu", "-ploidy", "unphased_diploid",
                                 "-posWmin1", str(u[5]), "-posWmax1", str(u[40]),
                                 "-posWmin2", str(u[20]), "-posWmax2", str(u[75]),
                                 "-r2limit", "0.05"],
        "unphased_samples": ["-input", "data/vu", "-ploidy", "unphased_diploid",
                             "-sampleList", "data/samples_vu.txt", "-r2limit", "0.02"],
        "unphased_two_inputs": ["-input", "data/vu", "-input2", "data/vu2", "-ploidy",
                                "unphased_diploid", "-posWmin1", str(u[0]), "-posWmax1",
                                str(u[79]), "-posWmin2", str(P["vu2"][0]),
                                "-posWmax2", str(P["vu2"][49]), "-r2limit", "0.03"],
        "unphased_list": ["-inputList", "list.csv", "-ploidy", "unphased_diploid",
                          "-r2limit", "0.05"],
        # the same list run largest task first: the shuffles follow the order
        # the list smallest task first: -sorted alone leaves the order (tasks
        # are fetched by id), -sorted -competing runs the largest first
        "unphased_list_sorted": ["-inputList", "list_sorted.csv", "-ploidy",
                                 "unphased_diploid", "-r2limit", "0.05", "-sorted"],
        "unphased_list_sorted_competing": ["-inputList", "list_sorted.csv", "-ploidy",
                                           "unphased_diploid", "-r2limit", "0.05", "-sorted",
                                           "-competing"],
    }


def list_text(P):
    u, u2 = P["vu"], P["vu2"]
    return "\n".join([
        f"data/vu,{u[2]},{u[50]},data/vu,{u[2]},{u[50]}",
        f"data/vu,{u[10]},{u[70]},data/vu2,{u2[3]},{u2[40]}",
        f"data/vu,{u[0]},{u[30]},data/vu,{u[40]},{u[79]}",
    ]) + "\n"


def sorted_list_text(lst):
    """The -inputList lines in reverse (smallest task first, so that -sorted
    reorders them), then a copy of one line (an exact tie: list order), a
    short window over the same chunks as a longer one (a tie in chunk count:
    the longer first) and a task on the second input alone."""
    lines = list(reversed(lst.strip().split("\n")))
    first = lines[0].split(",")
    short = ",".join([first[0], first[1], str(int(first[1]) + 40), first[3], first[4],
                      str(int(first[4]) + 40)])
    vu2 = [l for l in lst.strip().split("\n") if "vu2" in l][0].split(",")
    lines += [lines[-1], short, ",".join([vu2[3], vu2[4], vu2[5], vu2[3], vu2[4], vu2[5]])]
    return "\n".join(lines) + "\n"


def run_reference(name, args, lst):
    out = os.path.join(OUT, "expected", name)
    shutil.rmtree(out, ignore_errors=True)
    os.makedirs(out)
    work = os.path.join(OUT, "_run")
    shutil.rmtree(work, ignore_errors=True)
    os.makedirs(os.path.join(work, "out"))
    os.symlink(os.path.join(OUT, "data"), os.path.join(work, "data"))
    if "-inputList" in args:
        with open(os.path.join(work, "list.csv"), "w") as f:
            f.write(lst)
        with open(os.path.join(work, "list_sorted.csv"), "w") as f:
            f.write(sorted_list_text(lst))
    cmd = [REF_BIN] + args + ["-output", "out/report.txt", "-threads", "1"]
    r = subprocess.run(cmd, cwd=work, capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        raise SystemExit(f"{name}: reference exited {r.returncode}\n{r.stderr[-2000:]}")
    for fn in sorted(os.listdir(os.path.join(work, "out"))):
        shutil.copy(os.path.join(work, "out", fn), os.path.join(out, fn))
    shutil.copy(os.path.join(work, "input_report.txt"), os.path.join(out, "input_report.txt"))
    shutil.rmtree(work)


def main():
    if not os.path.exists(REF_BIN):
        raise SystemExit("build the reference CLI first: make -C oracle ref-cli")
    P = build()
    lst = list_text(P)
    with open(os.path.join(OUT, "list.csv"), "w") as f:
        f.write(lst)
    with open(os.path.join(OUT, "list_sorted.csv"), "w") as f:
        f.write(sorted_list_text(lst))
    cs = cases(P)
    for name, args in cs.items():
        run_reference(name, args, lst)
        print(name, sorted(os.listdir(os.path.join(OUT, "expected", name))), flush=True)
    with open(os.path.join(OUT, "cases.json"), "w") as f:
        json.dump({"cases": cs}, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
