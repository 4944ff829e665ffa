"""Write the qLD-compute fixtures on the reference's own test data under
tests/golden/covid/.

The reference's smoke script (test_run:29-30,48-56,77-92) parses
test_dataset/vcf/test_covid.vcf.gz with qLD-parse-VCF (-chrom 1 -size 1) and
runs qLD-compute on the result with -r2limit 0.6 -ploidy haploid.  This
script does the same with the reference's own builds (``make -C oracle
ref-parser ref-cli``):

1. the parser's chunk directory -> ``covid/data/covid_19/`` (22,554 haploid
   samples, 6,328 sites in 18 chunks, as the parser wrote them);
2. a sample list of every fourth name among the ``#CHROM`` line's names
   that occur once (the data repeats some names, and a listed name that
   matches two columns trips the reference's assertion in makeValidList,
   correlate_IO.c:287) -> ``covid/data/samples_covid.txt``;
3. for each case the report and ``input_report.txt`` of the reference
   correlator's CPU build, run without -mdf -> ``covid/expected/<case>/``;
4. the smoke script's MDF step: qLD-parse-2MDF with its default options
   (``make -C oracle ref-2mdf``; rows of the ploidy-2 width, 705 words for
   22,554 samples) into a scratch directory, the SHA-256 of every file it
   writes -> ``cases.json`` (``mdf``), and the reference correlator's -mdf
   reports on it -> ``covid/expected/<case>/`` for the ``mdf_cases``.  The
   8 MB MDF directory itself is not kept: the tests rebuild it with the
   VCF reader and MDF writer here and check it against these hashes.

Only this container has the reference checkout; everything written is data.
Usage: python tests/golden/make_covid_golden.py
"""
from __future__ import annotations

import collections
import gzip
import hashlib
import json
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "covid")
REF_VCF = "/root/reference/test_dataset/vcf/test_covid.vcf.gz"
REF_PARSER = os.path.join(ROOT, "oracle", "_ref", "qLD-parse-VCF")
REF_CLI = os.path.join(ROOT, "oracle", "_ref", "qLD-compute")
REF_2MDF = os.path.join(ROOT, "oracle", "_ref", "qLD-parse-2MDF")

CASES = {
    # test_run's own compute flags
    "test_run": ["-input", "data/covid_19", "-r2limit", "0.6", "-ploidy", "haploid"],
    "r2limit_0.1": ["-input", "data/covid_19", "-r2limit", "0.1", "-ploidy", "haploid"],
    "two_windows": ["-input", "data/covid_19", "-ploidy", "haploid",
                    "-posWmin1", "3000", "-posWmax1", "12000",
                    "-posWmin2", "9000", "-posWmax2", "25000", "-r2limit", "0.05"],
    "samples": ["-input", "data/covid_19", "-ploidy", "haploid",
                "-sampleList", "data/samples_covid.txt", "-r2limit", "0.3"],
}


# -inputList over the MDF directory: same-window triangles, a disjoint and an
# overlapping rectangle, and a line the reference rejects (posWmin > posWmax)
LIST_MDF = "\n".join([
    "mdf/covid_19,1000,9000,mdf/covid_19,1000,9000",
    "mdf/covid_19,2000,6000,mdf/covid_19,15000,22000",
    "mdf/covid_19,9000,8000,mdf/covid_19,100,200",
    "mdf/covid_19,20000,27000,mdf/covid_19,20000,27000",
    "mdf/covid_19,500,3000,mdf/covid_19,2500,8000",
]) + "\n"

MDF_CASES = {
    # test_run's compute step on its own MDF directory
    "test_run_mdf": ["-input", "mdf/covid_19", "-r2limit", "0.6", "-ploidy", "haploid", "-mdf"],
    "input_list_mdf": ["-inputList", "list_mdf.csv", "-ploidy", "haploid", "-mdf",
                       "-r2limit", "0.2"],
    "two_windows_mdf": ["-input", "mdf/covid_19", "-ploidy", "haploid", "-mdf",
                        "-posWmin1", "3000", "-posWmax1", "12000",
                        "-posWmin2", "9000", "-posWmax2", "25000", "-r2limit", "0.05"],
}


def parse():
    data = os.path.join(OUT, "data")
    shutil.rmtree(data, ignore_errors=True)
    os.makedirs(data)
    r = subprocess.run([REF_PARSER, "-input", REF_VCF, "-output",
                        os.path.join(data, "covid_19"), "-chrom", "1", "-size", "1"],
                       capture_output=True, text=True, timeout=300)
    # the parser ends with a non-zero status after writing every chunk
    if not os.path.exists(os.path.join(data, "covid_19", "1_header.vcf.gz")):
        raise SystemExit(f"parser failed ({r.returncode}):\n{r.stdout[-2000:]}{r.stderr[-2000:]}")
    with gzip.open(os.path.join(data, "covid_19", "1_header.vcf.gz"), "rt") as f:
        chrom = next(line for line in f if line.startswith("#CHROM"))
    names = chrom.rstrip("\n").split("\t")[9:]
    seen = collections.Counter(names)
    once = [n for n in names if seen[n] == 1]
    with open(os.path.join(data, "samples_covid.txt"), "w") as f:
        f.write("\n".join(once[::4]) + "\n")


def make_mdf(scratch):
    """qLD-parse-2MDF as test_run:59-65 runs it; returns {file: sha256}."""
    shutil.rmtree(scratch, ignore_errors=True)
    os.makedirs(scratch)
    mdf = os.path.join(scratch, "covid_19")
    r = subprocess.run([REF_2MDF, "-input", os.path.join(OUT, "data", "covid_19"),
                        "-output", mdf], capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise SystemExit(f"qLD-parse-2MDF exited {r.returncode}\n{r.stderr[-2000:]}")
    return {fn: hashlib.sha256(open(os.path.join(mdf, fn), "rb").read()).hexdigest()
            for fn in sorted(os.listdir(mdf))}


def run_reference(name, args, mdf_scratch=None):
    out = os.path.join(OUT, "expected", name)
    shutil.rmtree(out, ignore_errors=True)
    os.makedirs(out)
    work = os.path.join(OUT, "_run")
    shutil.rmtree(work, ignore_errors=True)
    os.makedirs(os.path.join(work, "out"))
    os.symlink(os.path.join(OUT, "data"), os.path.join(work, "data"))
    if mdf_scratch:
        os.symlink(mdf_scratch, os.path.join(work, "mdf"))
    if "-inputList" in args:
        shutil.copy(os.path.join(OUT, "list_mdf.csv"), os.path.join(work, "list_mdf.csv"))
    cmd = [REF_CLI] + args + ["-output", "out/report.txt", "-threads", "1"]
    r = subprocess.run(cmd, cwd=work, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise SystemExit(f"{name}: reference exited {r.returncode}\n{r.stderr[-2000:]}")
    for fn in sorted(os.listdir(os.path.join(work, "out"))):
        shutil.copy(os.path.join(work, "out", fn), os.path.join(out, fn))
    shutil.copy(os.path.join(work, "input_report.txt"), os.path.join(out, "input_report.txt"))
    shutil.rmtree(work)


def main():
    for b in (REF_PARSER, REF_CLI, REF_2MDF):
        if not os.path.exists(b):
            raise SystemExit("build the reference tools first: "
                             "make -C oracle ref-parser ref-cli ref-2mdf")
    parse()
    for name, args in CASES.items():
        run_reference(name, args)
        print(name, sorted(os.listdir(os.path.join(OUT, "expected", name))), flush=True)
    scratch = os.path.join(OUT, "_mdf")
    hashes = make_mdf(scratch)
    with open(os.path.join(OUT, "list_mdf.csv"), "w") as f:
        f.write(LIST_MDF)
    for name, args in MDF_CASES.items():
        run_reference(name, args, scratch)
        print(name, sorted(os.listdir(os.path.join(OUT, "expected", name))), flush=True)
    shutil.rmtree(scratch)
    with open(os.path.join(OUT, "cases.json"), "w") as f:
        json.dump({"cases": CASES, "mdf_cases": MDF_CASES,
                   "mdf": {"dir": "covid_19", "words": 705, "sha256": hashes}}, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
