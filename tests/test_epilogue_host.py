"""The REFERENCE r^2 epilogue the kernels compile (qld_amd/csrc/ld_epilogue.h),
built for the host and compared bit for bit with the oracle's
qo_r2_reference: every (a, b, c) for N <= 64 and about 9 million random
triples at the benchmark's sample counts, the largest accepted one and
random N.  The kernels compute each pair without a division and fall back to
the reference's division near float midpoints; this pins that the two agree
everywhere they were tried.  CPU only."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("epi") / "epilogue_check")
    oracle_dir = os.path.join(ROOT, "oracle")
    if not os.path.exists(os.path.join(oracle_dir, "liboracle.so")):
        subprocess.run(["make", "-s", "-C", oracle_dir, "liboracle.so"], check=True)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-I" + os.path.join(ROOT, "qld_amd", "csrc"), "-o", exe,
                    os.path.join(HERE, "epilogue_check.cpp"), "-L" + oracle_dir, "-loracle",
                    "-Wl,-rpath," + oracle_dir], check=True)
    return exe


def _tallies(out):
    got = {}
    for line in out.splitlines():
        w = line.split()
        got[w[0]] = {w[k]: int(w[k + 1]) for k in range(1, len(w) - 1, 2)}
    return got


def test_reference_epilogue_bit_exact_on_host(checker):
    r = subprocess.run([checker, "200"], capture_output=True, text=True,
                       env={**os.environ, "OMP_NUM_THREADS": "1"})
    assert r.returncode == 0, r.stdout + r.stderr
    t = _tallies(r.stdout)
    assert t["exhaustive"]["pairs"] == 814384
    assert t["random"]["pairs"] > 9_000_000
    for part in t.values():
        assert part["mismatches"] == 0
        # the fallback is taken wherever the fast value would be wrong
        assert 0 < part["needed"] <= part["slow"]
    # and rarely otherwise: well under 1% of the pairs
    assert t["random"]["slow"] < 0.01 * t["random"]["pairs"]
