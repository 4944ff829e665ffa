"""CPU oracle for the LD hot path -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg
as the checker.  The product libraries never import, link or call it.

* ``liboracle.so``      -- the C restatement (qld_oracle.c), built by
                           ``make -C oracle``.
* ``_ref/libqldref.so`` -- the reference's own C sources compiled where they
                           lie (``make -C oracle ref``); present only where
                           the reference checkout exists.  Used to pin the
                           restatement and to write tests/golden fixtures.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libqldref.so")

_o = None
_r = None
_libc = None


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def lib() -> C.CDLL:
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_LIB)
        vp = C.c_void_p
        L.qo_compress_snip_x64.restype = C.c_uint32
        L.qo_compress_snip_x64.argtypes = [C.c_char_p, C.c_int, vp]
        L.qo_counts_naive.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int, vp]
        L.qo_counts_blocked.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int, C.c_int,
                                        vp, C.c_int]
        for f in ("qo_r2_reference", "qo_r2_exact", "qo_r2_fp64"):
            getattr(L, f).argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int,
                                      C.c_int, vp]
        L.qo_counts_blocked_isa.restype = C.c_int
        L.qo_counts_blocked_isa.argtypes = []
        L.qo_fprint_f.restype = C.c_int
        L.qo_fprint_f.argtypes = [C.c_char_p, C.c_float, C.c_int]
        L.qo_write_results.argtypes = [vp, vp, vp, vp, C.c_int, vp, vp, vp,
                                       C.c_int, C.c_int, vp, C.c_float, C.c_int,
                                       C.c_int]
        _o = L
    return _o


def libc() -> C.CDLL:
    global _libc
    if _libc is None:
        _libc = C.CDLL("libc.so.6")
        _libc.fopen.restype = C.c_void_p
        _libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
        _libc.fclose.argtypes = [C.c_void_p]
        _libc.fflush.argtypes = [C.c_void_p]
    return _libc


# ---- restatement -----------------------------------------------------------

def pack_chars(chars: bytes) -> tuple[np.ndarray, int]:
    n = len(chars)
    row = np.zeros(max(1, (n + 63) // 64), dtype=np.uint64)
    bc = lib().qo_compress_snip_x64(chars, n, _p(row))
    return row, int(bc)


def counts_naive(A, B=None):
    A = np.ascontiguousarray(A, dtype=np.uint64)
    B = A if B is None else np.ascontiguousarray(B, dtype=np.uint64)
    m, W = A.shape
    n = B.shape[0]
    out = np.zeros((n, m), dtype=np.uint32)
    lib().qo_counts_naive(_p(A), m, _p(B), n, W, _p(out))
    return out


def counts_blocked_isa() -> int:
    """512 when counts_blocked runs its AVX-512 VPOPCNTQ micro-kernel here,
    64 for the scalar popcnt one."""
    return int(lib().qo_counts_blocked_isa())


def counts_blocked(A, B=None, tri=None, nthreads=0, out=None):
    A = np.ascontiguousarray(A, dtype=np.uint64)
    same = B is None
    B = A if same else np.ascontiguousarray(B, dtype=np.uint64)
    if tri is None:
        tri = same
    m, W = A.shape
    n = B.shape[0]
    if out is None:
        out = np.zeros((n, m), dtype=np.uint32)
    lib().qo_counts_blocked(_p(A), m, _p(B), n, W, 1 if tri else 0, _p(out), nthreads)
    return out


def _r2(fname, bcA, bcB, counts, snp_size, tri, dtype):
    bcA = np.ascontiguousarray(bcA, dtype=np.uint32)
    bcB = np.ascontiguousarray(bcB, dtype=np.uint32)
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    n, m = counts.shape
    out = np.zeros((n, m), dtype=dtype)
    getattr(lib(), fname)(_p(bcA), _p(bcB), _p(counts), m, n, int(snp_size),
                          1 if tri else 0, _p(out))
    return out


def r2_reference(bcA, bcB, counts, snp_size, tri=False):
    return _r2("qo_r2_reference", bcA, bcB, counts, snp_size, tri, np.float32)


def r2_exact(bcA, bcB, counts, snp_size, tri=False):
    return _r2("qo_r2_exact", bcA, bcB, counts, snp_size, tri, np.float32)


def r2_fp64(bcA, bcB, counts, snp_size, tri=False):
    return _r2("qo_r2_fp64", bcA, bcB, counts, snp_size, tri, np.float64)


def fprint_f(val: float, precision: int = 9) -> str:
    buf = C.create_string_buffer(64)
    n = lib().qo_fprint_f(buf, C.c_float(val), precision)
    return buf.raw[:n].decode()


def _id_array(ids, count):
    if ids is None:
        return None
    arr = (C.c_char_p * count)(*[s.encode() if isinstance(s, str) else s for s in ids])
    return C.cast(arr, C.c_void_p), arr  # keep arr alive via the tuple


# ctypes needs the char** itself; unwrap the keep-alive tuple at call sites.
def _ids_ptr(t):
    return None if t is None else t[0]


def write_results_text(path, posA, idA, bcA, posB, idB, bcB, snp_size, results,
                       r2limit, posWset2, header=True) -> str:
    """Report text as writeResults produces it; returns the file content."""
    posA = np.ascontiguousarray(posA, dtype=np.uint32)
    posB = np.ascontiguousarray(posB, dtype=np.uint32)
    bcA = np.ascontiguousarray(bcA, dtype=np.uint32)
    bcB = np.ascontiguousarray(bcB, dtype=np.uint32)
    results = np.ascontiguousarray(results, dtype=np.float32)
    n, m = results.shape
    ida = _id_array(idA, m)
    idb = _id_array(idB, n)
    fp = libc().fopen(str(path).encode(), b"w")
    lib().qo_write_results(fp, _p(posA), _ids_ptr(ida), _p(bcA), m, _p(posB),
                           _ids_ptr(idb), _p(bcB), n, int(snp_size), _p(results),
                           C.c_float(r2limit), int(posWset2), 1 if header else 0)
    libc().fclose(fp)
    with open(path) as f:
        return f.read()


# ---- the reference's own sources (only where the checkout exists) ---------

def ref_available() -> bool:
    return os.path.exists(REF_LIB)


class _Stats(C.Structure):
    _fields_ = [("job_counter", C.c_int), ("init_time", C.c_double),
                ("mlt_time", C.c_double), ("gemm_time", C.c_double),
                ("ld_time", C.c_double), ("write_time", C.c_double),
                ("total_lds", C.c_longlong)]


class _ThreadData(C.Structure):
    # threadData_t, src/correlator/include/pthreads.h:72-85
    _fields_ = [("threadID", C.c_int), ("threadTOTAL", C.c_int),
                ("threadOP", C.c_int), ("threadLog", C.c_void_p),
                ("threadStats", _Stats), ("r2limit", C.c_float),
                ("ploidy", C.c_int), ("gpu", C.c_int), ("blis", C.c_int),
                ("mdf", C.c_int), ("compQ", C.c_int), ("task_count", C.c_int)]


class _TableX64(C.Structure):
    # table_x64, src/correlator/include/correlate_MEM.h:66-75
    _fields_ = [("SNPtable", C.c_void_p), ("POStable", C.c_void_p),
                ("BCtable", C.c_void_p), ("IDtable", C.c_void_p),
                ("tableSize", C.c_uint), ("tableIndex", C.c_uint),
                ("compSize", C.c_uint), ("compIndex", C.c_uint)]


def ref() -> C.CDLL:
    global _r
    if _r is None:
        if not ref_available():
            raise RuntimeError(f"{REF_LIB} missing: run `make -C oracle ref`")
        _r = C.CDLL(REF_LIB)
    return _r


def ref_compress_snip(chars: bytes) -> tuple[np.ndarray, int]:
    """compressSnip_x64 of the reference on one SNP."""
    n = len(chars)
    W = max(1, (n + 63) // 64)
    words = np.zeros(W + 1, dtype=np.uint64)   # it may store one word past
    bc = np.zeros(1, dtype=np.uint32)
    t = _TableX64(words.ctypes.data, None, bc.ctypes.data, None, 1, 0, W, 0)
    R = ref()
    R.compressSnip_x64.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
    R.compressSnip_x64(C.byref(t), chars, n)
    return words[:W].copy(), int(bc[0])


def ref_counts(A, B=None):
    """Counts from the reference's own blocked popcount GEMM (gemm_cpu.c:25-341):
    mlt() transpose of A, then gemm() into a zeroed C, as correlate() runs it
    when BLIS is not selected."""
    A = np.ascontiguousarray(A, dtype=np.uint64)
    B = A if B is None else np.ascontiguousarray(B, dtype=np.uint64)
    m, k = A.shape
    n = B.shape[0]
    R = ref()
    At = np.zeros(m * k, dtype=np.uint64)
    R.mlt.argtypes = [C.c_uint, C.c_uint, C.c_void_p, C.c_void_p]
    R.mlt(m, k, _p(At), _p(A))
    Cm = np.zeros(n * m, dtype=np.uint64)
    ac = np.zeros(12 * 72 * 256, dtype=np.uint64)
    bcp = np.zeros(256 * 4032, dtype=np.uint64)
    R.gemm.argtypes = [C.c_uint, C.c_uint, C.c_uint, C.c_uint64, C.c_void_p,
                       C.c_uint, C.c_void_p, C.c_uint, C.c_void_p, C.c_uint,
                       C.c_void_p, C.c_void_p]
    R.gemm(m, n, k, 1, _p(At), m, _p(B), k, _p(Cm), m, _p(ac), _p(bcp))
    return Cm.reshape(n, m)


def ref_r2(bcA, bcB, counts_u64, snp_size):
    """get_pairwise_ld_score of the reference (full n x m square)."""
    bcA = np.ascontiguousarray(bcA, dtype=np.uint32)
    bcB = np.ascontiguousarray(bcB, dtype=np.uint32)
    Cm = np.ascontiguousarray(counts_u64, dtype=np.uint64)
    n, m = Cm.shape
    res = np.zeros(n * m, dtype=np.float32)
    resp = C.c_void_p(res.ctypes.data)
    R = ref()
    R.get_pairwise_ld_score.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_int, C.c_int, C.c_int, C.c_void_p]
    R.get_pairwise_ld_score(_p(bcA), _p(bcB), _p(Cm), m, n, int(snp_size),
                            C.byref(resp))
    return res.reshape(n, m)


def ref_fprint_f(val: float, path: str, precision: int = 9) -> str:
    R = ref()
    R.fprint_f.argtypes = [C.c_void_p, C.c_float, C.c_int]
    fp = libc().fopen(path.encode(), b"w")
    R.fprint_f(fp, C.c_float(val), precision)
    libc().fclose(fp)
    with open(path) as f:
        return f.read()


def ref_correlate_report(path, A, posA, idA, bcA, B, posB, idB, bcB,
                         snp_size, r2limit, posWset2) -> str:
    """The reference's full correlate() (gemm_cpu.c:420-623, non-BLIS gemm),
    writing its report to `path`; returns the text."""
    A = np.ascontiguousarray(A, dtype=np.uint64)
    B = np.ascontiguousarray(B, dtype=np.uint64)
    m, k = A.shape
    n = B.shape[0]
    posA = np.ascontiguousarray(posA, dtype=np.uint32)
    posB = np.ascontiguousarray(posB, dtype=np.uint32)
    bcA = np.ascontiguousarray(bcA, dtype=np.uint32)
    bcB = np.ascontiguousarray(bcB, dtype=np.uint32)
    ida = _id_array(idA, m)
    idb = _id_array(idB, n)
    log = libc().fopen(b"/dev/null", b"w")
    td = _ThreadData()
    td.threadLog = log
    td.r2limit = r2limit
    td.blis = 0
    td.gpu = 0
    fp = libc().fopen(str(path).encode(), b"w")
    R = ref()
    vp = C.c_void_p
    R.correlate.argtypes = [vp, vp, vp, vp, vp, vp, C.c_int, vp, vp, vp, vp,
                            C.c_int, C.c_int, C.c_int, C.c_int]
    R.correlate(C.byref(td), fp, _p(A), _p(posA), _ids_ptr(ida), _p(bcA), m,
                _p(B), _p(posB), _ids_ptr(idb), _p(bcB), n, k, int(snp_size),
                int(posWset2))
    libc().fclose(fp)
    libc().fclose(log)
    with open(path) as f:
        return f.read()
