"""Write the qLD-compute CLI fixtures under tests/golden/cli/.

1. Small seeded MDF directories (``cli/data/<id>/``): a gzip VCF header
   ``<id>_header.vcf.gz`` whose line after ``#CHROM`` is
   ``id snipsPerFile snipSize totalSnips posMin posMax`` (the parser writes
   it so, parse_vcf.c:620-634) and chunk files
   ``<id>_<first>_<last>_<posMin>_<posMax>.mdf`` in the MDF tool's text
   layout (process_data.c:447-460): ``##fileformat=MDFv0.1``,
   ``##SNPs:<k>``, ``#CHROM POS ID BC compSNP_0..``, rows
   ``chrom pos ID BC w0 w1 ..`` (packed words in decimal, MSB-first bits as
   compressSnip_x64, snip_processing.c:783-815), no trailing newline.
2. For each case below, the report (and ``input_report.txt``) written by the
   reference correlator's own CPU build, compiled from its sources by
   ``make -C oracle ref-cli`` (oracle/_ref/qLD-compute), run here with
   ``-threads 1`` -> ``cli/expected/<case>/``.

Only this container has the reference checkout; everything written is data.
Usage: python tests/golden/make_cli_golden.py
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
CLI = os.path.join(HERE, "cli")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "qLD-compute")


def pack(bits: np.ndarray) -> np.ndarray:
    """0/1 rows -> uint64 words, sample s at bit 63 - s % 64 of word s // 64."""
    rows, n = bits.shape
    words = (n + 63) // 64
    padded = np.zeros((rows, words * 64), dtype=np.uint8)
    padded[:, :n] = bits
    by = np.packbits(padded, axis=1, bitorder="big")          # 8 bytes per word
    return by.reshape(rows, words, 8).view(">u8").reshape(rows, words).astype(np.uint64)


def haplotypes(rng, rows, n, copy=0.85):
    out = np.zeros((rows, n), dtype=np.uint8)
    for s in range(rows):
        p = rng.uniform(0.03, 0.5)
        fresh = (rng.random(n) < p).astype(np.uint8)
        if s == 0 or rng.random() < 0.12:
            out[s] = fresh
        else:
            keep = rng.random(n) < copy
            out[s] = np.where(keep, out[s - 1], fresh)
    return out


def mdf_text(chrom, pos, ids, bc, words):
    """One chunk file in the MDF tool's layout."""
    k = words.shape[1]
    parts = [f"##fileformat=MDFv0.1\n##SNPs:{len(pos)}\n#CHROM\tPOS\tID\tBC"]
    parts.append("".join(f"\tcompSNP_{i}" for i in range(k)))
    for r in range(len(pos)):
        parts.append(f"\n{chrom}\t{int(pos[r])}\t{ids[r]}\t{int(bc[r])}")
        parts.append("".join(f"\t{int(w)}" for w in words[r]))
    return "".join(parts)


def write_dataset(name, bits, pos, snips_per_file, samples, snip_size, ids=None):
    d = os.path.join(CLI, "data", name)
    os.makedirs(d, exist_ok=True)
    rows = bits.shape[0]
    words = pack(bits)
    bc = bits.sum(axis=1).astype(np.uint32)
    if ids is None:
        ids = [f"rs{name}{i}" for i in range(rows)]
    hdr = ("##fileformat=VCFv4.2\n##source=qld_amd CLI fixture\n"
           "#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t"
           + "\t".join(f"S{i}" for i in range(samples)) + "\n"
           + f"{name} {snips_per_file} {snip_size} {rows} {int(pos[0])} {int(pos[-1])}\n")
    with gzip.GzipFile(os.path.join(d, f"{name}_header.vcf.gz"), "wb", mtime=0) as f:
        f.write(hdr.encode())
    for first in range(0, rows, snips_per_file):
        last = min(first + snips_per_file, rows)
        sl = slice(first, last)
        fn = f"{name}_{first + 1}_{last}_{int(pos[first])}_{int(pos[last - 1])}.mdf"
        with open(os.path.join(d, fn), "w") as f:
            f.write(mdf_text(name, pos[sl], ids[first:last], bc[sl], words[sl]))
    return d


def positions(rng, rows, start, gaps_at=()):
    steps = rng.integers(1, 40, size=rows)
    for g in gaps_at:
        steps[g] += 500
    pos = start + np.cumsum(steps) - steps[0]
    return pos.astype(np.int64)


def build_data():
    rng = np.random.default_rng(20261203)
    shutil.rmtree(os.path.join(CLI, "data"), ignore_errors=True)
    # chrA: 96 SNPs x 150 haploid samples (3 words), chunks of 25; special rows
    N = 150
    a = haplotypes(rng, 96, N)
    a[5] = 0                                  # monomorphic (bc = 0 -> r2 0)
    a[17] = 1                                 # all ones (bc = N -> NaN, never printed)
    a[30] = a[29]                             # duplicate (r2 = 1)
    a[41] = 1 - a[40]                         # complement (r2 = 1)
    pos_a = positions(rng, 96, 1000, gaps_at=(50, 75))
    write_dataset("chrA", a, pos_a, 25, N, N)
    # chrB: 70 SNPs, same samples, chunks of 30
    b = haplotypes(rng, 70, N)
    b[3] = a[60]                              # shared haplotype with chrA
    pos_b = positions(rng, 70, 5000)
    write_dataset("chrB", b, pos_b, 30, N, N)
    # chrC: 60 SNPs whose positions span chrA's whole range (two-input runs
    # without windows search chrC with chrA's range)
    c = haplotypes(rng, 60, N)
    pos_c = (900 + np.arange(60) * 80).astype(np.int64)
    assert pos_c[0] < pos_a[0] and pos_c[-1] > pos_a[-1]
    write_dataset("chrC", c, pos_c, 20, N, N)
    # dip: 40 phased diploid samples -> 80 haplotypes (2 words), one chunk
    d = haplotypes(rng, 50, 80)
    pos_d = positions(rng, 50, 200)
    write_dataset("dip", d, pos_d, 50, 40, 80)
    # zp: positions from 0, so the default window [0, posMax] takes each
    # chunk's #CHROM line (POS reads as 0) as a row
    z = haplotypes(rng, 30, 70)
    pos_z = positions(rng, 30, 0)
    write_dataset("zp", z, pos_z, 12, 70, 70)
    return {"chrA": pos_a, "chrB": pos_b, "chrC": pos_c, "dip": pos_d, "zp": pos_z}


def cases(P):
    a, b = P["chrA"], P["chrB"]
    return {
        "single_all": ["-input", "data/chrA", "-ploidy", "haploid"],
        "single_win": ["-input", "data/chrA", "-posWmin1", str(a[10]), "-posWmax1", str(a[70]),
                       "-r2limit", "0.1"],
        "single_win_gap": ["-input", "data/chrA", "-posWmin1", str(a[49] + 100),
                           "-posWmax1", str(a[75] - 100), "-r2limit", "0.05"],
        "single_two_windows": ["-input", "data/chrA", "-posWmin1", str(a[3]),
                               "-posWmax1", str(a[40]), "-posWmin2", str(a[20]),
                               "-posWmax2", str(a[90])],
        "single_win2_only": ["-input", "data/chrA", "-posWmin2", str(a[10]),
                             "-posWmax2", str(a[30]), "-r2limit", "0.3"],
        "two_inputs_windows": ["-input", "data/chrA", "-input2", "data/chrB",
                               "-posWmin1", str(a[2]), "-posWmax1", str(a[80]),
                               "-posWmin2", str(b[1]), "-posWmax2", str(b[65]),
                               "-r2limit", "0.02"],
        "two_inputs_nowin": ["-input", "data/chrA", "-input2", "data/chrC"],
        "two_inputs_win1_only": ["-input", "data/chrA", "-input2", "data/chrB",
                                 "-posWmin1", str(a[0]), "-posWmax1", str(a[95])],
        "two_inputs_win1_equal": ["-input", "data/chrA", "-input2", "data/chrC",
                                  "-posWmin1", str(a[4]), "-posWmax1", str(a[90])],
        "same_input_twice": ["-input", "data/chrA", "-input2", "data/chrA/",
                             "-posWmin1", str(a[10]), "-posWmax1", str(a[60]),
                             "-posWmin2", str(a[30]), "-posWmax2", str(a[95])],
        "r2limit0": ["-input", "data/chrA", "-posWmin1", str(a[0]), "-posWmax1", str(a[35]),
                     "-r2limit", "0"],
        "window_clamped": ["-input", "data/chrA", "-posWmin1", "-5", "-posWmax1", "99999999",
                           "-r2limit", "0.5"],
        "diploid": ["-input", "data/dip", "-ploidy", "phased_diploid", "-r2limit", "0.05"],
        "zero_pos": ["-input", "data/zp", "-r2limit", "0"],
        "input_list": ["-inputList", "list.csv"],
    }


# Runs on which the reference never returns (its findFiles rescans the
# directory forever when no chunk continues the window); qLD-compute must
# exit with an error instead.  Not run on the reference.
def error_cases(P):
    a = P["chrA"]
    return {
        "two_disjoint_inputs_nowin": ["-input", "data/chrA", "-input2", "data/chrB",
                                      "-ploidy", "haploid"],
        "missing_header": ["-input", "data/empty", "-ploidy", "haploid"],
        "no_ploidy": ["-input", "data/chrA"],
    }


def input_list(P):
    a, b = P["chrA"], P["chrB"]
    return "\n".join([
        f"data/chrA,{a[5]},{a[60]},data/chrA,{a[5]},{a[60]}",      # single region
        f"data/chrA,{a[0]},{a[30]},data/chrA,{a[50]},{a[95]}",     # two windows
        "not,a,valid,line",                                        # bad structure
        f"data/chrA,{a[40]},{a[10]},data/chrA,{a[5]},{a[60]}",     # posWmin >= posWmax
        f"data/nope,{a[0]},{a[30]},data/chrA,{a[5]},{a[60]}",      # missing input
        f"data/chrA,{a[20]},{a[85]},data/chrB,{b[0]},{b[69]}",     # two inputs
    ]) + "\n"


def run_reference(name, args, list_text):
    out = os.path.join(CLI, "expected", name)
    shutil.rmtree(out, ignore_errors=True)
    os.makedirs(out)
    work = os.path.join(CLI, "_run")
    shutil.rmtree(work, ignore_errors=True)
    os.makedirs(work)
    os.symlink(os.path.join(CLI, "data"), os.path.join(work, "data"))
    if name == "input_list":
        with open(os.path.join(work, "list.csv"), "w") as f:
            f.write(list_text)
    if "-ploidy" not in args:
        args = args + ["-ploidy", "haploid"]
    cmd = [REF_BIN] + args + ["-output", "out/report.txt", "-mdf", "-threads", "1"]
    os.makedirs(os.path.join(work, "out"))
    r = subprocess.run(cmd, cwd=work, capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        raise SystemExit(f"{name}: reference exited {r.returncode}\n{r.stderr}")
    for fn in sorted(os.listdir(os.path.join(work, "out"))):
        shutil.copy(os.path.join(work, "out", fn), os.path.join(out, fn))
    shutil.copy(os.path.join(work, "input_report.txt"), os.path.join(out, "input_report.txt"))
    shutil.rmtree(work)
    return args


def main():
    if not os.path.exists(REF_BIN):
        raise SystemExit("build the reference CLI first: make -C oracle ref-cli")
    P = build_data()
    lst = input_list(P)
    with open(os.path.join(CLI, "list.csv"), "w") as f:
        f.write(lst)
    manifest = {}
    for name, args in cases(P).items():
        manifest[name] = run_reference(name, args, lst)
        files = sorted(os.listdir(os.path.join(CLI, "expected", name)))
        print(name, files, flush=True)
    os.makedirs(os.path.join(CLI, "data", "empty"), exist_ok=True)
    with open(os.path.join(CLI, "data", "empty", "README"), "w") as f:
        f.write("an input directory without a header file\n")
    with open(os.path.join(CLI, "cases.json"), "w") as f:
        json.dump({"cases": manifest, "errors": error_cases(P)}, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
