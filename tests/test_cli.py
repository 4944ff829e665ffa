"""qLD-compute (qld_amd/lib/qLD-compute) against the reference correlator.

The golden reports under tests/golden/cli/expected/ were written by the
reference's own CPU build (tests/golden/make_cli_golden.py) on the MDF
directories under tests/golden/cli/data/.

CPU: the CLI loads each case's tables (-dump-tables, no GPU) and the oracle
computes counts, r^2 and the report text from them; the text must equal the
reference's byte for byte, which pins header discovery, chunk selection,
window clamping and row loading.  input_report.txt must match as well.
GPU: the CLI runs for real, on every engine, and its report files must equal
the reference's byte for byte.
"""
import json
import os
import shutil
import struct
import subprocess
import sys

import numpy as np
import pytest

import oracle
import qld_amd
from conftest import GOLDEN, ROOT, unpack_bits

CLI_DIR = os.path.join(GOLDEN, "cli")
CLI_BIN = os.path.join(ROOT, "qld_amd", "lib", "qLD-compute")
MANIFEST = json.load(open(os.path.join(CLI_DIR, "cases.json")))
CASES = sorted(MANIFEST["cases"])


def _workdir(tmp_path, name):
    os.symlink(os.path.join(CLI_DIR, "data"), tmp_path / "data")
    os.makedirs(tmp_path / "out")
    if name == "input_list":
        shutil.copy(os.path.join(CLI_DIR, "list.csv"), tmp_path / "list.csv")
    return tmp_path


def _run(tmp_path, args, extra=(), timeout=120):
    cmd = [CLI_BIN] + list(args) + ["-output", "out/report.txt", "-mdf"] + list(extra)
    return subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True,
                          timeout=timeout)


def _read_tables(path):
    buf = open(path, "rb").read()
    assert buf[:4] == b"QLDT"
    single, wset2, snp_size = struct.unpack_from("<3i", buf, 4)
    off = 16

    def table():
        nonlocal off
        rows, words = struct.unpack_from("<2i", buf, off)
        off += 8
        pos = np.frombuffer(buf, np.uint32, rows, off); off += 4 * rows
        bc = np.frombuffer(buf, np.uint32, rows, off); off += 4 * rows
        snp = np.frombuffer(buf, np.uint64, rows * words, off).reshape(rows, words)
        off += 8 * rows * words
        ids = []
        for _ in range(rows):
            e = buf.index(b"\0", off)
            ids.append(buf[off:e].decode())
            off = e + 1
        return dict(pos=pos, bc=bc, snp=snp, id=ids, words=words)

    A = table()
    B = A if single else table()
    assert off == len(buf)
    return single, wset2, snp_size, A, B


def _oracle_report(path, A, B, snp_size, wset2, r2limit):
    """Counts (exact integer matrix product of the unpacked bits), the
    reference's float r^2 and writeResults text, all from the oracle."""
    m, n = len(A["pos"]), len(B["pos"])
    if m and n:
        UA = unpack_bits(A["snp"], A["words"] * 64)
        UB = unpack_bits(B["snp"], B["words"] * 64)
        counts = (UB @ UA.T).astype(np.uint32)
        r2 = oracle.r2_reference(A["bc"], B["bc"], counts, snp_size)
    else:
        r2 = np.zeros((n, m), np.float32)
    return oracle.write_results_text(path, A["pos"], A["id"], A["bc"], B["pos"], B["id"],
                                     B["bc"], snp_size, r2, r2limit, wset2)


def _r2limit(args):
    return float(args[args.index("-r2limit") + 1]) if "-r2limit" in args else 0.2


def _expected(name):
    d = os.path.join(CLI_DIR, "expected", name)
    return {f: open(os.path.join(d, f), "rb").read() for f in sorted(os.listdir(d))}


@pytest.mark.parametrize("name", CASES)
def test_cli_tables_reproduce_reference_report(name, tmp_path):
    args = MANIFEST["cases"][name]
    wd = _workdir(tmp_path, name)
    r = _run(wd, args, ["-dump-tables"])
    assert r.returncode == 0, r.stderr
    exp = _expected(name)
    assert open(wd / "input_report.txt", "rb").read() == exp["input_report.txt"]
    dumps = sorted(f for f in os.listdir(wd / "out") if f.endswith(".tables"))
    reports = sorted(f for f in exp if f != "input_report.txt")
    assert [d[:-len(".tables")] for d in dumps] == reports
    for d, rep in zip(dumps, reports):
        single, wset2, snp_size, A, B = _read_tables(wd / "out" / d)
        text = _oracle_report(wd / "out" / (rep + ".oracle"), A, B, snp_size, wset2,
                              _r2limit(args))
        assert text.encode() == exp[rep], (name, rep)


def test_cli_zero_pos_takes_chrom_lines_as_rows(tmp_path):
    """With a window reaching POS 0 every chunk's #CHROM line becomes a row
    (POS 0, ID "ID", BC 0, words 0), as the reference reads it."""
    args = MANIFEST["cases"]["zero_pos"]
    wd = _workdir(tmp_path, "zero_pos")
    assert _run(wd, args, ["-dump-tables"]).returncode == 0
    _, _, _, A, _ = _read_tables(wd / "out" / "report.txt.tables")
    hdr = qld_amd.mdf_read_header(os.path.join(CLI_DIR, "data", "zp"))
    chunks = -(-hdr.total_snips // hdr.snips_per_file)
    fake = [i for i, s in enumerate(A["id"]) if s == "ID"]
    assert len(fake) == chunks
    assert all(A["pos"][i] == 0 and A["bc"][i] == 0 and not A["snp"][i].any() for i in fake)
    assert len(A["pos"]) == hdr.total_snips + chunks


@pytest.mark.parametrize("name", sorted(MANIFEST["errors"]))
def test_cli_error_cases_exit_instead_of_hanging(name, tmp_path):
    wd = _workdir(tmp_path, name)
    r = _run(wd, MANIFEST["errors"][name], ["-dump-tables"], timeout=60)
    assert r.returncode != 0
    assert "ERROR" in r.stderr or "ploidy" in r.stderr


def test_cli_batch_cells_argument_is_checked(tmp_path):
    """-batch-cells takes 4 .. 2^40 r^2 cells; anything else stops the
    command before any input is read."""
    wd = _workdir(tmp_path, "single_all")
    for bad in (["-batch-cells", "2"], ["-batch-cells", str(1 << 41)], ["-batch-cells"]):
        r = subprocess.run([CLI_BIN, "-input", "data/chrA", "-ploidy", "haploid", "-mdf",
                            "-output", "out/report.txt"] + bad,
                           cwd=wd, capture_output=True, text=True, timeout=60)
        assert r.returncode != 0 and "-batch-cells" in r.stderr, bad
        assert os.listdir(wd / "out") == []


def test_cli_vcf_mode_on_an_mdf_directory_and_unknown_flags(tmp_path):
    """Without -mdf the chunks are .vcf.gz files: an MDF directory has none,
    which is an error here (the reference rescans forever)."""
    wd = _workdir(tmp_path, "x")
    r = subprocess.run([CLI_BIN, "-input", "data/chrA", "-ploidy", "haploid", "-output",
                        "out/r.txt", "-dump-tables"], cwd=wd, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "no chunk" in r.stderr
    r = _run(wd, ["-input", "data/chrA", "-ploidy", "haploid", "-bogus"])
    assert r.returncode == 1 and "-bogus is not a valid" in r.stderr


def test_mdf_writer_reproduces_fixture_files():
    """Rows read back from every fixture chunk and written again with
    qld_mdf_write give the same bytes (the MDF tool's layout)."""
    for name in ("chrA", "chrB", "chrC", "dip"):
        d = os.path.join(CLI_DIR, "data", name)
        for fn in sorted(os.listdir(d)):
            if not fn.endswith(".mdf"):
                continue
            path = os.path.join(d, fn)
            words = int(open(path).read().split("\n", 3)[2].split("\t")[-1].split("_")[1]) + 1
            t = qld_amd.mdf_read_table([path], 1, 2**31 - 1, words)
            out = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"qld_mdf_{os.getpid()}.mdf")
            qld_amd.mdf_write(out, name, t["snp"], t["pos"], t["id"], t["bc"])
            same = open(out, "rb").read() == open(path, "rb").read()
            os.unlink(out)
            assert same, path


def test_mdf_rows_match_bit_counts():
    """Loaded words are the packed rows: popcount equals the BC column and the
    padding bits past the sample count are zero."""
    for name, n in (("chrA", 150), ("chrB", 150), ("dip", 80)):
        d = os.path.join(CLI_DIR, "data", name)
        files = sorted(os.path.join(d, f) for f in os.listdir(d) if f.endswith(".mdf"))
        t = qld_amd.mdf_read_table(files, 1, 2**31 - 1, (n + 63) // 64)
        bits = unpack_bits(t["snp"], t["snp"].shape[1] * 64)
        assert (bits.sum(axis=1) == t["bc"]).all()
        assert not bits[:, n:].any()


def test_mdf_parallel_read_is_deterministic():
    d = os.path.join(CLI_DIR, "data", "chrA")
    h = qld_amd.mdf_read_header(d)
    files = qld_amd.mdf_find_files(d, "chrA", 25, 96, h.pos_min, h.pos_max)
    assert len(files) == 4
    t1 = qld_amd.mdf_read_table(files, h.pos_min, h.pos_max, 3, nthreads=1)
    t8 = qld_amd.mdf_read_table(files, h.pos_min, h.pos_max, 3, nthreads=8)
    assert len(t1["pos"]) == 96
    assert (t1["snp"] == t8["snp"]).all() and t1["id"] == t8["id"]
    assert (t1["pos"] == t8["pos"]).all() and (t1["bc"] == t8["bc"]).all()


def test_mdf_find_files_window_past_last_chunk_is_an_error():
    """A window ending past the last chunk makes the reference rescan the
    directory forever; here it is an error."""
    d = os.path.join(CLI_DIR, "data", "chrA")
    with pytest.raises(qld_amd.QldError, match="continues"):
        qld_amd.mdf_find_files(d, "chrA", 25, 96, 1000, 10**6)


def _chrB_chunk():
    src = os.path.join(CLI_DIR, "data", "chrB", "chrB_61_70_6181_6393.mdf")
    if not os.path.exists(src):
        src = sorted(os.path.join(CLI_DIR, "data", "chrB", f)
                     for f in os.listdir(os.path.join(CLI_DIR, "data", "chrB"))
                     if f.endswith(".mdf"))[-1]
    return src


def test_mdf_row_with_too_few_words_is_an_error(tmp_path):
    text = open(_chrB_chunk()).read().split("\n")
    text[4] = text[4].rsplit("\t", 1)[0]          # drop one word of the second row
    bad = tmp_path / "bad.mdf"
    bad.write_text("\n".join(text))
    with pytest.raises(qld_amd.QldError, match="fewer packed words"):
        qld_amd.mdf_read_table([str(bad)], 1, 2**31 - 1, 3)


def test_mdf_rows_wider_than_the_ploidy_keep_their_first_words(tmp_path):
    """The MDF tool writes ploidy-2 rows unless told otherwise, and the
    reference's smoke run reads them as haploid (test_run:59-92): its row
    loop keeps each row's first words (correlate_IOx64.c:102-115), so wider
    rows -- zero or not past the header's width -- load as their first
    words."""
    src = _chrB_chunk()
    want = qld_amd.mdf_read_table([src], 1, 2**31 - 1, 3)
    text = open(src).read().split("\n")
    for k in range(3, len(text)):
        text[k] += "\t0\t0\t0" if k % 2 else "\t18446744073709551615\t7"
    wide = tmp_path / "wide.mdf"
    wide.write_text("\n".join(text))
    got = qld_amd.mdf_read_table([str(wide)], 1, 2**31 - 1, 3)
    for key in ("snp", "pos", "bc"):
        assert (got[key] == want[key]).all(), key
    assert got["id"] == want["id"]


@pytest.mark.gpu
@pytest.mark.parametrize("engine", ["auto", "popcount", "mfma_fp4", "mfma_i8"])
def test_cli_reports_equal_reference_on_gpu(engine, tmp_path):
    for name in CASES:
        wd = tmp_path / f"{name}_{engine}"
        os.makedirs(wd)
        _workdir(wd, name)
        print(f"[cli] {name} engine={engine}", flush=True)
        r = _run(wd, MANIFEST["cases"][name], ["-engine", engine])
        assert r.returncode == 0, (name, r.stderr)
        exp = _expected(name)
        got = {f: open(wd / "out" / f, "rb").read() for f in sorted(os.listdir(wd / "out"))}
        got["input_report.txt"] = open(wd / "input_report.txt", "rb").read()
        assert sorted(got) == sorted(exp), name
        for f in exp:
            assert got[f] == exp[f], (name, f, engine)


@pytest.mark.gpu
def test_cli_batched_input_list_equals_single_tasks(tmp_path):
    """-inputList runs its tasks as one batched plan; every report must equal
    the report of the same line run as a single task (which the golden cases
    pin to the reference)."""
    rng = np.random.default_rng(99)
    a = qld_amd.mdf_read_header(os.path.join(CLI_DIR, "data", "chrA"))
    b = qld_amd.mdf_read_header(os.path.join(CLI_DIR, "data", "chrB"))
    pa = np.arange(a.pos_min, a.pos_max)
    pb = np.arange(b.pos_min, b.pos_max)
    lines = []
    for k in range(40):
        w1 = np.sort(rng.choice(pa, 2, replace=False))
        kind = k % 3
        if kind == 0:
            lines.append(("data/chrA", w1, "data/chrA", w1))
        elif kind == 1:
            lines.append(("data/chrA", w1, "data/chrA", np.sort(rng.choice(pa, 2, replace=False))))
        else:
            lines.append(("data/chrA", w1, "data/chrB", np.sort(rng.choice(pb, 2, replace=False))))
    os.symlink(os.path.join(CLI_DIR, "data"), tmp_path / "data")
    os.makedirs(tmp_path / "out")
    with open(tmp_path / "list.csv", "w") as f:
        for i1, w1, i2, w2 in lines:
            f.write(f"{i1},{w1[0]},{w1[1]},{i2},{w2[0]},{w2[1]}\n")
    print("[cli] batched list of", len(lines), "tasks", flush=True)
    r = subprocess.run([CLI_BIN, "-inputList", "list.csv", "-ploidy", "haploid", "-mdf",
                        "-r2limit", "0.05", "-output", "out/lst.txt"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    for n, (i1, w1, i2, w2) in enumerate(lines, start=1):
        print(f"[cli] single task {n}", flush=True)
        r = subprocess.run([CLI_BIN, "-input", i1, "-posWmin1", str(w1[0]), "-posWmax1",
                            str(w1[1]), "-input2", i2, "-posWmin2", str(w2[0]), "-posWmax2",
                            str(w2[1]), "-ploidy", "haploid", "-mdf", "-r2limit", "0.05",
                            "-output", f"out/single_{n}.txt"], cwd=tmp_path,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        got = open(tmp_path / "out" / f"lst_{n}.txt", "rb").read()
        want = open(tmp_path / "out" / f"single_{n}.txt", "rb").read()
        assert got == want, (n, lines[n - 1])


def test_mdf_reader_cache_matches_one_shot_reads():
    """Windows read through one caching reader (files read and rows parsed
    once) equal independent one-shot reads, in any order, with repeats."""
    d = os.path.join(CLI_DIR, "data", "chrA")
    h = qld_amd.mdf_read_header(d)
    rng = np.random.default_rng(5)
    reader = qld_amd.MdfReader()
    for _ in range(25):
        lo, hi = np.sort(rng.integers(h.pos_min, h.pos_max + 1, 2))
        files = qld_amd.mdf_find_files(d, "chrA", 25, 96, int(lo), int(hi))
        for wset in (1, 0):
            t1 = qld_amd.mdf_read_table(files, int(lo), int(hi), 3, pos_wset=wset)
            t2 = reader.table(files, int(lo), int(hi), 3, pos_wset=wset)
            assert (t1["snp"] == t2["snp"]).all() and t1["id"] == t2["id"]
            assert (t1["pos"] == t2["pos"]).all() and (t1["bc"] == t2["bc"]).all()
    with pytest.raises(qld_amd.QldError, match="fewer packed words"):
        reader.table(files, h.pos_min, h.pos_max, 4)     # wider packing: re-parsed, rejected
    narrow = reader.table(files, h.pos_min, h.pos_max, 2)   # narrower: each row's first words
    full = qld_amd.mdf_read_table(files, h.pos_min, h.pos_max, 3)
    assert (narrow["snp"] == full["snp"][:, :2]).all()
    t = reader.table(files, int(lo), int(hi), 3)
    assert len(t["pos"]) == len(qld_amd.mdf_read_table(files, int(lo), int(hi), 3)["pos"])
    reader.close()


# ---- VCF chunk input (tests/golden/vcf/, written by make_vcf_golden.py) ----

VCF_DIR = os.path.join(GOLDEN, "vcf")
VCF_MANIFEST = json.load(open(os.path.join(VCF_DIR, "cases.json")))
VCF_CASES = sorted(VCF_MANIFEST["cases"])


def _vcf_workdir(tmp_path):
    os.symlink(os.path.join(VCF_DIR, "data"), tmp_path / "data")
    os.makedirs(tmp_path / "out")
    shutil.copy(os.path.join(VCF_DIR, "list.csv"), tmp_path / "list.csv")
    shutil.copy(os.path.join(VCF_DIR, "list_sorted.csv"), tmp_path / "list_sorted.csv")
    return tmp_path


def _vcf_expected(name):
    d = os.path.join(VCF_DIR, "expected", name)
    return {f: open(os.path.join(d, f), "rb").read() for f in sorted(os.listdir(d))}


@pytest.mark.parametrize("name", VCF_CASES)
def test_cli_vcf_tables_reproduce_reference_report(name, tmp_path):
    """VCF chunks (haploid, phased and unphased diploid, sample lists, two
    inputs, -inputList): the CLI's tables through the oracle give the
    reference's report byte for byte -- row filters, allele coding, the
    rand() shuffles of unphased genotypes and BC included."""
    args = VCF_MANIFEST["cases"][name]
    wd = _vcf_workdir(tmp_path)
    r = subprocess.run([CLI_BIN] + args + ["-output", "out/report.txt", "-dump-tables"],
                       cwd=wd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    exp = _vcf_expected(name)
    assert open(wd / "input_report.txt", "rb").read() == exp["input_report.txt"]
    dumps = sorted(f for f in os.listdir(wd / "out") if f.endswith(".tables"))
    reports = sorted(f for f in exp if f != "input_report.txt")
    assert [d[:-len(".tables")] for d in dumps] == reports
    for d, rep in zip(dumps, reports):
        single, wset2, snp_size, A, B = _read_tables(wd / "out" / d)
        text = _oracle_report(wd / "out" / (rep + ".oracle"), A, B, snp_size, wset2,
                              _r2limit(args))
        assert text.encode() == exp[rep], (name, rep)


@pytest.mark.gpu
@pytest.mark.parametrize("cells", [4, 3000, 1 << 29])
def test_cli_vcf_input_list_batch_limits(cells, tmp_path):
    """-inputList over unphased VCF chunks with batch limits that send every
    task alone, mix lone tasks and batches, or batch them all: each task's
    tables are read once, in list order (the unphased shuffles draw from one
    rand() stream), so the reports equal the reference's whatever the split."""
    name = "unphased_list"
    wd = _vcf_workdir(tmp_path)
    print(f"[cli-vcf] {name} batch-cells={cells}", flush=True)
    r = subprocess.run([CLI_BIN] + VCF_MANIFEST["cases"][name] +
                       ["-output", "out/report.txt", "-batch-cells", str(cells)],
                       cwd=wd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    exp = _vcf_expected(name)
    got = {f: open(wd / "out" / f, "rb").read() for f in sorted(os.listdir(wd / "out"))}
    got["input_report.txt"] = open(wd / "input_report.txt", "rb").read()
    assert sorted(got) == sorted(exp)
    for f in exp:
        assert got[f] == exp[f], (f, cells)


def test_cli_seed_restarts_the_unphased_shuffles(tmp_path):
    """-seed S restarts the reader's rand() stream as srand(S) restarts the
    reference's: S = 1 is glibc's default seed (the golden tables), another
    seed shuffles the unphased genotypes differently, the same way on every
    run.  (The reference's own parser seeds rand() and then rejects the
    flag, correlate_our_kernel.c:387-392,433-435, so no reference report
    pins a seeded run.)"""
    name = "unphased"
    args = VCF_MANIFEST["cases"][name]

    def tables(seed, tag):
        wd = tmp_path / tag
        os.makedirs(wd)
        _vcf_workdir(wd)
        extra = [] if seed is None else ["-seed", str(seed)]
        r = subprocess.run([CLI_BIN] + args + ["-output", "out/report.txt", "-dump-tables"] +
                           extra, cwd=wd, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        return open(wd / "out" / "report.txt.tables", "rb").read()

    default = tables(None, "default")
    assert tables(1, "seed1") == default
    seven = tables(7, "seed7")
    assert seven != default
    assert tables(7, "seed7b") == seven


def test_vcf_reader_rand_stream_is_glibc_default():
    """The reader's private generator reproduces the C library's unseeded
    rand() sequence (which the reference's shuffles draw from)."""
    import ctypes as C
    libc = C.CDLL(None)
    libc.rand.restype = C.c_int
    # a fresh process has the default state; check through a child
    out = subprocess.run([sys.executable, "-c",
                          "import ctypes as C; l=C.CDLL(None); "
                          "print(' '.join(str(l.rand()) for _ in range(8)))"],
                         capture_output=True, text=True, timeout=60).stdout.split()
    assert [int(x) for x in out][:3] == [1804289383, 846930886, 1681692777]


@pytest.mark.gpu
def test_cli_vcf_reports_equal_reference_on_gpu(tmp_path):
    for name in VCF_CASES:
        wd = tmp_path / name
        os.makedirs(wd)
        _vcf_workdir(wd)
        print(f"[cli-vcf] {name}", flush=True)
        r = subprocess.run([CLI_BIN] + VCF_MANIFEST["cases"][name] + ["-output",
                                                                      "out/report.txt"],
                           cwd=wd, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (name, r.stderr)
        exp = _vcf_expected(name)
        got = {f: open(wd / "out" / f, "rb").read() for f in sorted(os.listdir(wd / "out"))}
        got["input_report.txt"] = open(wd / "input_report.txt", "rb").read()
        assert sorted(got) == sorted(exp), name
        for f in exp:
            assert got[f] == exp[f], (name, f)


# ---- the reference's own test data (tests/golden/covid/, make_covid_golden.py) ----

COVID_DIR = os.path.join(GOLDEN, "covid")
COVID_MANIFEST = json.load(open(os.path.join(COVID_DIR, "cases.json")))
COVID_CASES = sorted(COVID_MANIFEST["cases"])


def _covid_workdir(tmp_path):
    os.symlink(os.path.join(COVID_DIR, "data"), tmp_path / "data")
    os.makedirs(tmp_path / "out")
    return tmp_path


def _covid_expected(name):
    d = os.path.join(COVID_DIR, "expected", name)
    return {f: open(os.path.join(d, f), "rb").read() for f in sorted(os.listdir(d))}


@pytest.mark.parametrize("name", COVID_CASES)
def test_cli_covid_tables_reproduce_reference_report(name, tmp_path):
    """The reference's own test VCF (test_dataset/vcf/test_covid.vcf.gz:
    22,554 haploid samples, 6,328 sites) as its parser chunks it, under the
    compute flags of its smoke script (test_run:77-92) and three variations
    (r2limit 0.1, two windows, a sample list): the CLI's tables through the
    oracle give the reference's report byte for byte."""
    args = COVID_MANIFEST["cases"][name]
    wd = _covid_workdir(tmp_path)
    r = subprocess.run([CLI_BIN] + args + ["-output", "out/report.txt", "-dump-tables"],
                       cwd=wd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    exp = _covid_expected(name)
    assert open(wd / "input_report.txt", "rb").read() == exp["input_report.txt"]
    assert sorted(os.listdir(wd / "out")) == ["report.txt.tables"]
    single, wset2, snp_size, A, B = _read_tables(wd / "out" / "report.txt.tables")
    assert snp_size == 22554 and len(A["pos"]) > 1000
    counts = oracle.counts_blocked(A["snp"], B["snp"], tri=False)
    r2 = oracle.r2_reference(A["bc"], B["bc"], counts, snp_size)
    text = oracle.write_results_text(str(wd / "out" / "report.txt.oracle"), A["pos"], A["id"],
                                     A["bc"], B["pos"], B["id"], B["bc"], snp_size, r2,
                                     _r2limit(args), wset2)
    assert text.encode() == exp["report.txt"], name


def test_cli_sample_list_matching_a_repeated_name_is_an_error(tmp_path):
    """test_covid's header names one sample twice.  A sample list naming it
    selects two columns for one name, which the reference's makeValidList
    asserts against (correlate_IO.c:287-288); here the command fails with a
    message instead."""
    wd = _covid_workdir(tmp_path)
    (wd / "dup.txt").write_text("HCOV-19_SPAIN_20200326_EPI_ISL_436284\n"
                                "HCOV-19_SPAIN_20200324_EPI_ISL_436287\n")
    r = subprocess.run([CLI_BIN, "-input", "data/covid_19", "-ploidy", "haploid",
                        "-sampleList", "dup.txt", "-output", "out/report.txt",
                        "-dump-tables"], cwd=wd, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "2 names select 3 sample columns" in r.stderr + r.stdout
    assert os.listdir(wd / "out") == []


COVID_MDF_CASES = sorted(COVID_MANIFEST["mdf_cases"])


@pytest.fixture(scope="module")
def covid_mdf(tmp_path_factory):
    """The smoke script's MDF directory (test_run:59-65, qLD-parse-2MDF with
    its default ploidy-2 width), rebuilt here by the VCF reader and the MDF
    writer: every file must hash as the reference converter's did."""
    import hashlib
    m = COVID_MANIFEST["mdf"]
    root = tmp_path_factory.mktemp("covid_mdf")
    src = os.path.join(COVID_DIR, "data", m["dir"])
    dst = root / "mdf" / m["dir"]
    os.makedirs(dst)
    hdr = next(f for f in sorted(os.listdir(src)) if "header" in f)
    shutil.copy(os.path.join(src, hdr), dst / hdr)
    rd = qld_amd.VcfReader(os.path.join(src, hdr), 1)
    for fn in sorted(os.listdir(src)):
        if fn == hdr:
            continue
        t = rd.table([os.path.join(src, fn)], 0, 2**31 - 1, 22554)
        snp = np.zeros((len(t["pos"]), m["words"]), np.uint64)
        snp[:, :t["snp"].shape[1]] = t["snp"]
        qld_amd.mdf_write(str(dst / (fn[:-len(".vcf.gz")] + ".mdf")), "1", snp, t["pos"],
                          t["id"], t["bc"])
    rd.close()
    got = {fn: hashlib.sha256(open(dst / fn, "rb").read()).hexdigest()
           for fn in sorted(os.listdir(dst))}
    assert got == m["sha256"]
    return root


def _covid_mdf_workdir(tmp_path, covid_mdf):
    os.symlink(covid_mdf / "mdf", tmp_path / "mdf")
    os.makedirs(tmp_path / "out")
    shutil.copy(os.path.join(COVID_DIR, "list_mdf.csv"), tmp_path / "list_mdf.csv")
    return tmp_path


@pytest.mark.parametrize("name", COVID_MDF_CASES)
def test_cli_covid_mdf_tables_reproduce_reference_report(name, covid_mdf, tmp_path):
    """The smoke script's compute step on its MDF directory (-mdf -ploidy
    haploid over 705-word rows), two windows, and an -inputList of five
    lines (triangles, rectangles, a rejected line): the CLI's tables through
    the oracle give the reference's reports byte for byte."""
    args = COVID_MANIFEST["mdf_cases"][name]
    wd = _covid_mdf_workdir(tmp_path, covid_mdf)
    r = subprocess.run([CLI_BIN] + args + ["-output", "out/report.txt", "-dump-tables"],
                       cwd=wd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    exp = _covid_expected(name)
    assert open(wd / "input_report.txt", "rb").read() == exp["input_report.txt"]
    dumps = sorted(f for f in os.listdir(wd / "out") if f.endswith(".tables"))
    reports = sorted(f for f in exp if f != "input_report.txt")
    assert [d[:-len(".tables")] for d in dumps] == reports
    for d, rep in zip(dumps, reports):
        single, wset2, snp_size, A, B = _read_tables(wd / "out" / d)
        assert A["words"] == 353
        counts = oracle.counts_blocked(A["snp"], B["snp"], tri=False)
        r2 = oracle.r2_reference(A["bc"], B["bc"], counts, snp_size)
        text = oracle.write_results_text(str(wd / "out" / (rep + ".oracle")), A["pos"],
                                         A["id"], A["bc"], B["pos"], B["id"], B["bc"],
                                         snp_size, r2, _r2limit(args), wset2)
        assert text.encode() == exp[rep], (name, rep)


@pytest.mark.gpu
@pytest.mark.parametrize("name", COVID_MDF_CASES)
def test_cli_covid_mdf_reports_equal_reference_on_gpu(name, covid_mdf, tmp_path):
    wd = _covid_mdf_workdir(tmp_path, covid_mdf)
    print(f"[cli-covid-mdf] {name}", flush=True)
    r = subprocess.run([CLI_BIN] + COVID_MANIFEST["mdf_cases"][name] +
                       ["-output", "out/report.txt"],
                       cwd=wd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (name, r.stderr)
    exp = _covid_expected(name)
    got = {f: open(wd / "out" / f, "rb").read() for f in sorted(os.listdir(wd / "out"))}
    got["input_report.txt"] = open(wd / "input_report.txt", "rb").read()
    assert sorted(got) == sorted(exp), name
    for f in exp:
        assert got[f] == exp[f], (name, f)


@pytest.mark.gpu
@pytest.mark.parametrize("name,engine",
                         [(n, "auto") for n in COVID_CASES] +
                         [("test_run", e) for e in ("popcount", "mfma_fp4", "mfma_i8")])
def test_cli_covid_reports_equal_reference_on_gpu(name, engine, tmp_path):
    wd = _covid_workdir(tmp_path)
    print(f"[cli-covid] {name} engine={engine}", flush=True)
    r = subprocess.run([CLI_BIN] + COVID_MANIFEST["cases"][name] +
                       ["-output", "out/report.txt", "-engine", engine],
                       cwd=wd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (name, r.stderr)
    exp = _covid_expected(name)
    got = {f: open(wd / "out" / f, "rb").read() for f in sorted(os.listdir(wd / "out"))}
    got["input_report.txt"] = open(wd / "input_report.txt", "rb").read()
    assert sorted(got) == sorted(exp), name
    for f in exp:
        assert got[f] == exp[f], (name, f, engine)


def _write_many_mdf(tmp_path, nfiles, rows_per_file=4, words=3, seed=11):
    """nfiles chunk files of consecutive positions; returns their paths."""
    rng = np.random.default_rng(seed)
    d = tmp_path / "many"
    d.mkdir()
    paths = []
    for f in range(nfiles):
        snp = rng.integers(0, 2**63, (rows_per_file, words), dtype=np.uint64)
        pos = np.arange(rows_per_file, dtype=np.uint32) + 1000 + f * rows_per_file
        bc = np.array([sum(bin(int(w)).count("1") for w in r) for r in snp], dtype=np.uint32)
        p = str(d / f"c_{f * rows_per_file + 1}_{(f + 1) * rows_per_file}_{pos[0]}_{pos[-1]}.mdf")
        qld_amd.mdf_write(p, "c", snp, pos, [f"r{f}_{k}" for k in range(rows_per_file)], bc)
        paths.append(p)
    return paths


def test_mdf_reader_many_files_and_bounded_cache(tmp_path):
    """More than 64 and more than 128 distinct files through one reader, in
    one call and across calls (the registry grows while earlier files are in
    use), equal one-shot reads; under a byte limit the cache stays bounded
    by the limit plus the files of the latest call."""
    paths = _write_many_mdf(tmp_path, 200)
    one = qld_amd.mdf_read_table(paths, 0, 10**9, 3)
    reader = qld_amd.MdfReader()
    got = reader.table(paths, 0, 10**9, 3)
    assert (got["snp"] == one["snp"]).all() and got["id"] == one["id"]
    reader.close()
    per_file = None
    reader = qld_amd.MdfReader(limit=0)
    for a in list(range(0, 200, 35)) + list(range(0, 200, 35)):
        group = paths[a:a + 70]
        t = reader.table(group, 0, 10**9, 3)
        want = qld_amd.mdf_read_table(group, 0, 10**9, 3)
        assert (t["snp"] == want["snp"]).all() and (t["pos"] == want["pos"]).all()
        per_file = per_file or reader.cached_bytes // len(group)
        # limit 0: only the latest call's files stay
        assert reader.cached_bytes <= len(group) * per_file * 1.01
    reader.close()
    reader = qld_amd.MdfReader(limit=40 * per_file)
    for a in range(0, 200, 10):
        reader.table(paths[a:a + 10], 0, 10**9, 3)
        assert reader.cached_bytes <= 41 * per_file
    reader.close()


def test_mdf_reader_many_files_asan(tmp_path):
    """The same access pattern through the C API in a build of the MDF reader
    with AddressSanitizer: no invalid access, and the tables match a build
    without it."""
    import shutil
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    paths = _write_many_mdf(tmp_path, 150)
    src = [os.path.join(ROOT, "tests", "mdf_reader_many.c"),
           os.path.join(ROOT, "qld_amd", "host", "mdf.c")]
    outs = {}
    for name, flags in (("plain", ["-O1"]), ("asan", ["-O1", "-g", "-fsanitize=address",
                                                      "-fno-omit-frame-pointer"])):
        exe = str(tmp_path / f"mdf_many_{name}")
        r = subprocess.run([gcc, "-std=gnu11", "-fopenmp", *flags,
                            f"-I{os.path.join(ROOT, 'include')}", "-o", exe, *src, "-lz"],
                           capture_output=True, text=True)
        if r.returncode and name == "asan":
            pytest.skip("no AddressSanitizer runtime: " + r.stderr[-200:])
        assert r.returncode == 0, r.stderr
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0")
        r = subprocess.run([exe, "3", "20000", *paths], capture_output=True, text=True,
                           env=env, timeout=120)
        assert r.returncode == 0 and "ERROR: AddressSanitizer" not in r.stderr, r.stderr[-3000:]
        outs[name] = r.stdout.splitlines()
    assert outs["plain"] == outs["asan"]
    assert outs["plain"][0].startswith("all 600 ")


@pytest.mark.skipif(not os.path.isdir("/root/reference/src/correlator"),
                    reason="needs the reference checkout (build container only)")
def test_reference_gpu_binding_builds_and_keeps_cpu_path(tmp_path):
    """The binding a maintainer adds to the reference (oracle/ref_gpu_binding.c:
    gpu_init / gpu_release / correlate_gpu over the engine's C ABI, and the
    per-task dispatch of the CPU task driver) compiles and links against the
    reference's own correlator sources with no OpenCL left; without -gpu the
    bound command still writes the reference's golden report (the CPU
    threads keep correlate()), and with -gpu it reaches the engine's init."""
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref-gpu"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    exe = os.path.join(ROOT, "oracle", "_ref", "qLD-compute-gpu")
    syms = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True,
                          text=True).stdout
    names = {ln.split()[-1].split("@")[0] for ln in syms.splitlines() if ln.strip()}
    assert {"qld_correlate_report", "qld_gpu_init", "qld_gpu_release"} <= names
    assert not any(n.startswith("cl") and n[2:3].isupper() for n in names)   # no OpenCL
    os.symlink(os.path.join(CLI_DIR, "data"), tmp_path / "data")
    os.makedirs(tmp_path / "out")
    args = MANIFEST["cases"]["single_win"]
    r = subprocess.run([exe] + args + ["-mdf", "-threads", "1", "-output", "out/report.txt"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    want = open(os.path.join(CLI_DIR, "expected", "single_win", "report.txt"), "rb").read()
    assert open(tmp_path / "out" / "report.txt", "rb").read() == want
    r = subprocess.run([exe] + args + ["-mdf", "-threads", "1", "-gpu", "-output",
                                       "out/report_gpu.txt"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    if r.returncode != 0:       # no GPU in this container: the engine's init says so
        assert "GPU init failed" in r.stderr, r.stderr
