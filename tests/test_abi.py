"""C-ABI and host-side checks that need no GPU: every declared symbol is
exported, the host report writer matches the golden text, the synthetic
generator is deterministic and well formed, and the engine rejects bad
arguments before touching a device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
import qld_amd
from conftest import ROOT, golden_cases, load_case, unpack_bits


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(qld_\w+)\s*\(", src)))


@pytest.mark.parametrize("header,libpath", [("qld_gpu.h", qld_amd.HIP_LIB),
                                            ("qld_host.h", qld_amd.HOST_LIB)])
def test_library_exports_every_declared_symbol(header, libpath):
    names = _declared(header)
    assert names, header
    lib = C.CDLL(libpath, mode=C.RTLD_GLOBAL)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_table_covers_header():
    bound = {n for n, _, _ in qld_amd.GPU_SYMBOLS} | {n for n, _, _ in qld_amd.HOST_SYMBOLS}
    declared = set(_declared("qld_gpu.h")) | set(_declared("qld_host.h"))
    assert declared <= bound, declared - bound


def test_hip_library_has_gfx950_code():
    blob = open(qld_amd.HIP_LIB, "rb").read()
    assert b"gfx950" in blob


def test_format_float_matches_golden():
    z = np.load(os.path.join(ROOT, "tests", "golden", "fprint_kat.npz"))
    for v, t in zip(z["vals"], z["texts"]):
        assert qld_amd.format_float(float(v)) == str(t)


def _host_write(path, d, r2):
    lib = qld_amd.host_lib()
    libc = oracle.libc()
    m, n = len(d["posA"]), len(d["posB"])
    ida = (C.c_char_p * m)(*[s.encode() for s in d["idA"]])
    idb = (C.c_char_p * n)(*[s.encode() for s in d["idB"]])
    fp = libc.fopen(str(path).encode(), b"w")
    lib.qld_write_results_header(fp)
    p = lambda a: np.ascontiguousarray(a).ctypes.data_as(C.c_void_p)  # noqa: E731
    lines = lib.qld_write_results(fp, p(d["posA"]), C.cast(ida, C.c_void_p), p(d["bcA"]), m,
                                  p(d["posB"]), C.cast(idb, C.c_void_p), p(d["bcB"]), n,
                                  int(d["snp_size"]), p(r2.astype(np.float32)),
                                  C.c_float(float(d["r2limit"])), int(d["posWset2"]))
    libc.fclose(fp)
    return open(path).read(), lines


@pytest.mark.parametrize("path", golden_cases(), ids=lambda p: os.path.basename(p))
def test_host_writer_matches_reference_report(path, tmp_path):
    d = load_case(path)
    text, lines = _host_write(tmp_path / "o.txt", d, d["r2"])
    assert text == d["report"]
    assert lines == text.count("\n") - 1


def test_synth_deterministic_and_thread_independent():
    a1, b1 = qld_amd.synth_blocks(42, 300, 1000, nthreads=1)
    a4, b4 = qld_amd.synth_blocks(42, 300, 1000, nthreads=4)
    assert (a1 == a4).all() and (b1 == b4).all()
    a5, _ = qld_amd.synth_blocks(43, 300, 1000)
    assert not (a1 == a5).all()


@pytest.mark.parametrize("N", [1, 63, 64, 65, 16622, 100000])
def test_synth_layout(N):
    t, bc = qld_amd.synth_blocks(7, 40, N)
    W = (N + 63) // 64
    assert t.shape == (40, W)
    if N % 64:
        assert (t[:, -1] & np.uint64((1 << (64 - N % 64)) - 1) == 0).all()
    U = unpack_bits(t, N)
    assert (U.sum(axis=1) == bc).all()


def test_synth_statistics():
    """MAF within the requested range on average; adjacent SNPs correlated."""
    N = 16622
    t, bc = qld_amd.synth_blocks(20261205, 2000, N)
    f = bc / N
    assert 0.005 < f.min() and f.max() < 0.56
    assert abs(f.mean() - 0.255) < 0.03
    c = oracle.counts_naive(t[:200])
    r2 = oracle.r2_fp64(bc[:200], bc[:200], c, N)
    adj = np.array([r2[i, i + 1] for i in range(199)])
    far = np.array([r2[i, (i + 100) % 200] for i in range(199)])
    assert np.nanmean(adj) > 0.3 and np.nanmean(far) < 0.05


def test_synth_regions_copies():
    N = 2000
    A, bA, B, bB = qld_amd.synth_regions(5, 300, 400, N)
    c = oracle.counts_naive(A, B)
    r2 = oracle.r2_fp64(bA, bB, c, N)
    best = np.nanmax(r2, axis=1)          # per B SNP, best A partner
    frac = (best > 0.8).mean()
    assert 0.02 < frac < 0.10             # ~5% copied with 2% noise
    f = np.concatenate([bA, bB]) / N
    assert f.min() >= 0.001 and f.max() <= 0.53   # MAF clip [0.005, 0.5] + noise


def test_engine_init_rejects_bad_config():
    with pytest.raises(qld_amd.QldError) as e:
        qld_amd.Engine(engine=9)
    assert e.value.code == qld_amd.QLD_ERR_INVALID


def test_tile_index_matches_tile_list(tmp_path):
    """The closed-form position of a tile in a job's work list
    (job_tile_index, which the device tile-list kernel of the allocation-free
    entry point uses) equals its position in append_job_tiles' list for every
    tile grid up to 70 x 70, rectangles and triangles; and the per-wave block
    lists of plans (append_job_wave_units, both kernel geometries) cover every
    cell a job computes exactly once with no empty block, in row-band order
    inside each super-tile, and their reordering by work is a permutation
    with the most blocks first (a host program built from the same
    header)."""
    import subprocess
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "tile_order_check")
    r = subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O1", "-o", exe,
                        os.path.join(ROOT, "tests", "tile_order_check.cpp")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("ok 4970 544"), r.stdout + r.stderr
