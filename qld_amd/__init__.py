"""MI355X-native LD engine for qLD-compute (Python glue over the C ABI).

The product is the pair of native libraries built in ``qld_amd/lib``:
``libqld_hip.so`` (HIP kernels behind ``include/qld_gpu.h``) and
``libqld_host.so`` (C host side, ``include/qld_host.h``).  This module only
loads them with ctypes for tests, benchmarks and scripting; it contains no
compute path of its own, and every call fails loudly when the HIP library
is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(PKG, "lib")
HIP_LIB = os.path.join(LIBDIR, "libqld_hip.so")
HOST_LIB = os.path.join(LIBDIR, "libqld_host.so")

QLD_OK = 0
QLD_ERR_INVALID = -1
QLD_ERR_NOMEM = -2
QLD_ERR_DEVICE = -3
QLD_ERR_BUDGET = -4
QLD_ERR_NOTREADY = -5
QLD_ERR_NOTFOUND = -6

ENGINES = {"auto": 0, "popcount": 1, "mfma_i8": 2, "mfma_fp4": 3}
ENGINE_NAMES = {v: k for k, v in ENGINES.items()}
R2_MODES = {"reference": 0, "exact": 1}


class QldError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"qld error {code}: {msg}")
        self.code = code


class Cfg(C.Structure):
    _fields_ = [("device", C.c_int), ("engine", C.c_int), ("r2_mode", C.c_int),
                ("timing", C.c_int), ("hbm_budget", C.c_size_t)]


_u64p = C.POINTER(C.c_uint64)
_u32p = C.POINTER(C.c_uint32)
_f32p = C.POINTER(C.c_float)
_vp = C.c_void_p

# (name, restype, argtypes) of every symbol include/qld_gpu.h declares.
GPU_SYMBOLS = [
    ("qld_cfg_default", None, [C.POINTER(Cfg)]),
    ("qld_gpu_init", C.c_int, [C.POINTER(_vp), C.POINTER(Cfg)]),
    ("qld_gpu_release", None, [_vp]),
    ("qld_gpu_last_error", C.c_char_p, [_vp]),
    ("qld_gpu_correlate", C.c_int, [_vp, _vp, _vp, C.c_int, _vp, _vp, C.c_int,
                                    C.c_int, C.c_int, C.c_int, _vp, _vp]),
    ("qld_dev_alloc", C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    ("qld_dev_free", C.c_int, [_vp, _vp]),
    ("qld_h2d", C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    ("qld_d2h", C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    ("qld_dev_fill", C.c_int, [_vp, _vp, C.c_uint32, C.c_size_t]),
    ("qld_dev_copy", C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    ("qld_sync", C.c_int, [_vp]),
    ("qld_correlate_dev", C.c_int, [_vp, _vp, _vp, C.c_int, _vp, _vp, C.c_int,
                                    C.c_int, C.c_int, C.c_int, _vp, C.c_int64,
                                    _vp, C.c_int64]),
    ("qld_correlate_dev_noalloc", C.c_int, [_vp, _vp, _vp, C.c_int, _vp, _vp, C.c_int,
                                            C.c_int, C.c_int, C.c_int, _vp, C.c_int64,
                                            _vp, C.c_int64]),
    ("qld_last_kernel_ms", C.c_int, [_vp, _f32p]),
    ("qld_set_plan_schedule", C.c_int, [_vp, C.c_int]),
    ("qld_get_plan_schedule", C.c_int, [_vp]),
    ("qld_set_plan_shared", C.c_int, [_vp, C.c_int]),
    ("qld_get_plan_shared", C.c_int, [_vp]),
    ("qld_set_plan_graphs", C.c_int, [_vp, C.c_int]),
    ("qld_last_contract_ms", C.c_int, [_vp, _f32p]),
    ("qld_timer_start", C.c_int, [_vp]),
    ("qld_timer_stop", C.c_int, [_vp, _f32p]),
    ("qld_last_engine", C.c_int, [_vp]),
    ("qld_set_plain_tile", C.c_int, [_vp, C.c_int]),
    ("qld_set_ksplit", C.c_int, [_vp, C.c_int]),
    ("qld_set_timing", C.c_int, [_vp, C.c_int]),
    ("qld_last_ksplit", C.c_int, [_vp]),
    ("qld_set_fp4_operands", C.c_int, [_vp, C.c_int]),
    ("qld_table_create", C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    ("qld_table_create_dev", C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(_vp)]),
    ("qld_table_free", C.c_int, [_vp, _vp]),
    ("qld_table_rows", C.c_int, [_vp]),
    ("qld_table_bytes", C.c_size_t, [_vp]),
    ("qld_plan_windows", C.c_int, [_vp, _vp, _vp, C.c_int, _vp, C.c_int, C.POINTER(_vp)]),
    ("qld_plan_pairs", C.c_int, [_vp, _vp, _vp, C.c_int, _vp, C.c_int, C.POINTER(_vp)]),
    ("qld_plan_run", C.c_int, [_vp, _vp]),
    ("qld_plan_use_graph", C.c_int, [_vp, C.c_int]),
    ("qld_plan_stats", C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int)]),
    ("qld_plan_free", C.c_int, [_vp, _vp]),
    ("qld_server_create", C.c_int, [_vp, _vp, C.POINTER(_vp)]),
    ("qld_server_destroy", C.c_int, [_vp]),
    ("qld_server_submit", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, _vp,
                                    C.POINTER(C.c_int64)]),
    ("qld_server_wait", C.c_int, [_vp, C.c_int64, C.c_double, C.POINTER(C.c_double)]),
    ("qld_server_fetch", C.c_int, [_vp, C.c_int64, _vp]),
    ("qld_server_release", C.c_int, [_vp, C.c_int64]),
    ("qld_server_stats", C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64)]),
    ("qld_host_alloc", C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    ("qld_host_free", C.c_int, [_vp, _vp]),
    ("qld_set_tile_cutoff", C.c_int, [_vp, C.c_int]),
    ("qld_get_tile_cutoff", C.c_int, [_vp]),
    ("qld_autotune_windows", C.c_int, [_vp, _vp, _vp, C.c_int, _vp, C.c_int,
                                       C.POINTER(C.c_int), _f32p]),
    ("qld_correlate_hits", C.c_int, [_vp, _vp, _vp, C.c_int, _vp, _vp, C.c_int, C.c_int,
                                     C.c_int, C.c_int, C.c_float, C.c_int64, _vp, _vp]),
    ("qld_compact_dev", C.c_int, [_vp, _vp, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_int, C.c_float, _vp, _vp, _vp, _vp, C.c_uint64, _vp]),
]

REQ_KEEP = 0
REQ_COPY = 1
REQ_DISCARD = 2


class ServerCfg(C.Structure):
    _fields_ = [("out_budget", C.c_size_t), ("max_batch", C.c_int),
                ("streams", C.c_int), ("reserved", C.c_int)]


WINDOW_DTYPE = np.dtype([("start", "<i4"), ("size", "<i4"), ("out_offset", "<i8")])
PAIR_WINDOW_DTYPE = np.dtype([("a_start", "<i4"), ("a_size", "<i4"), ("b_start", "<i4"),
                              ("b_size", "<i4"), ("out_offset", "<i8"), ("tri", "<i4"),
                              ("reserved", "<i4")])
BATCH_PLAIN = 0
BATCH_PERSISTENT = 1
BATCH_SHARED = 2

HOST_SYMBOLS = [
    ("qld_format_float", C.c_int, [C.c_char_p, C.c_float, C.c_int]),
    ("qld_write_results_header", None, [_vp]),
    ("qld_write_results", C.c_long, [_vp, _vp, _vp, _vp, C.c_int, _vp, _vp, _vp,
                                     C.c_int, C.c_int, _vp, C.c_float, C.c_int]),
    ("qld_correlate_report", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int,
                                       _vp, _vp, _vp, _vp, C.c_int, C.c_int,
                                       C.c_int, C.c_int, C.c_float]),
    ("qld_write_hits", C.c_long, [_vp, _vp, _vp, _vp, C.c_int, _vp, _vp, _vp, C.c_int,
                                  C.c_int, C.c_int, C.c_int, _vp, _vp, _vp]),
    ("qld_words", C.c_int, [C.c_int]),
    ("qld_synth_blocks", None, [C.c_uint64, C.c_int, C.c_int, C.c_double,
                                C.c_double, C.c_double, C.c_double, _vp, _vp,
                                C.c_int]),
    ("qld_synth_regions", None, [C.c_uint64, C.c_int, C.c_int, C.c_int,
                                 C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_double, _vp, _vp, _vp, _vp,
                                 C.c_int]),
    ("qld_bitcounts", None, [_vp, C.c_int, C.c_int, _vp, C.c_int]),
    ("qld_host_last_error", C.c_char_p, []),
    ("qld_mdf_read_header", C.c_int, [C.c_char_p, _vp]),
    ("qld_mdf_find_files", C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int,
                                     C.c_int, _vp, _vp]),
    ("qld_mdf_free_files", None, [_vp, C.c_int]),
    ("qld_find_chunks", C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_char_p, _vp, _vp]),
    ("qld_vcf_reader_create", _vp, []),
    ("qld_vcf_reader_free", None, [_vp]),
    ("qld_vcf_reader_seed", None, [_vp, C.c_uint]),
    ("qld_vcf_last_error", C.c_char_p, []),
    ("qld_vcf_header_lines", C.c_int, [C.c_char_p, _vp, _vp]),
    ("qld_vcf_sample_mask", C.c_int, [C.c_char_p, C.c_char_p, C.c_int, _vp, _vp, _vp]),
    ("qld_vcf_read_table", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p,
                                     C.c_char_p, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp]),
    ("qld_mdf_read_table", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, _vp]),
    ("qld_snp_table_free", None, [_vp]),
    ("qld_mdf_reader_create", _vp, []),
    ("qld_mdf_reader_table", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_int, _vp]),
    ("qld_mdf_reader_free", None, [_vp]),
    ("qld_mdf_reader_set_limit", None, [_vp, C.c_size_t]),
    ("qld_mdf_reader_bytes", C.c_size_t, [_vp]),
    ("qld_mdf_write", C.c_int, [C.c_char_p, C.c_char_p, _vp, _vp, _vp, _vp, C.c_int,
                                C.c_int]),
]

PATH_MAX = 4096

_HITS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_uint64),
                       C.POINTER(C.c_uint32), C.POINTER(C.c_float))


class MdfHeader(C.Structure):
    """qld_mdf_header (include/qld_host.h)."""
    _fields_ = [("id", C.c_char * 256), ("header_path", C.c_char * PATH_MAX),
                ("snips_per_file", C.c_int), ("snip_size", C.c_int),
                ("total_snips", C.c_int), ("pos_min", C.c_int), ("pos_max", C.c_int),
                ("samples", C.c_int)]


class _SnpTable(C.Structure):
    _fields_ = [("rows", C.c_int), ("words", C.c_int), ("snp", C.c_void_p),
                ("pos", C.c_void_p), ("bc", C.c_void_p), ("id", C.c_void_p),
                ("idbuf", C.c_void_p)]

_hip = None
_host = None


def _bind(lib, table):
    for name, res, args in table:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def hip_lib() -> C.CDLL:
    """The HIP engine library; raises if it has not been built."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise QldError(QLD_ERR_DEVICE,
                           f"{HIP_LIB} is missing: run `python -m qld_amd.build` "
                           "(there is no CPU fallback)")
        lib = C.CDLL(HIP_LIB, mode=C.RTLD_GLOBAL)
        _bind(lib, GPU_SYMBOLS)
        _hip = lib
    return _hip


def host_lib() -> C.CDLL:
    global _host
    if _host is None:
        hip_lib()
        if not os.path.exists(HOST_LIB):
            raise QldError(QLD_ERR_DEVICE, f"{HOST_LIB} is missing: run "
                           "`python -m qld_amd.build`")
        lib = C.CDLL(HOST_LIB)
        _bind(lib, HOST_SYMBOLS)
        _host = lib
    return _host


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def words_for(n_samples: int) -> int:
    return (n_samples + 63) // 64


# ---- synthetic data (host C generator) ------------------------------------

def synth_blocks(seed: int, n_snps: int, n_samples: int, mean_block=50.0,
                 maf_lo=0.01, maf_hi=0.5, copy_prob=0.9, nthreads=0):
    """Haplotype-block Markov table: (table[n_snps, W] uint64, bc uint32)."""
    W = words_for(n_samples)
    table = np.empty((n_snps, W), dtype=np.uint64)
    bc = np.empty(n_snps, dtype=np.uint32)
    host_lib().qld_synth_blocks(seed, n_snps, n_samples, mean_block, maf_lo,
                                maf_hi, copy_prob, _ptr(table), _ptr(bc), nthreads)
    return table, bc


def synth_regions(seed: int, m_a: int, n_b: int, n_samples: int,
                  mean_block=50.0, maf_lo=0.005, maf_hi=0.5, copy_prob=0.9,
                  copy_frac=0.05, flip_prob=0.02, nthreads=0):
    W = words_for(n_samples)
    ta = np.empty((m_a, W), dtype=np.uint64)
    tb = np.empty((n_b, W), dtype=np.uint64)
    ba = np.empty(m_a, dtype=np.uint32)
    bb = np.empty(n_b, dtype=np.uint32)
    host_lib().qld_synth_regions(seed, m_a, n_b, n_samples, mean_block, maf_lo,
                                 maf_hi, copy_prob, copy_frac, flip_prob,
                                 _ptr(ta), _ptr(ba), _ptr(tb), _ptr(bb), nthreads)
    return ta, ba, tb, bb


# ---- MDF directories (host C reader/writer) --------------------------------

def _host_check(rc: int, what: str) -> None:
    if rc != QLD_OK:
        msg = host_lib().qld_host_last_error() or b""
        raise QldError(rc, f"{what}: {msg.decode(errors='replace')}")


def mdf_read_header(directory: str) -> MdfHeader:
    h = MdfHeader()
    _host_check(host_lib().qld_mdf_read_header(directory.encode(), C.byref(h)),
                f"header of {directory}")
    return h


def mdf_find_files(directory: str, hid: str, snips_per_file: int, total_snips: int,
                   pos_wmin: int, pos_wmax: int) -> list[str]:
    files = C.POINTER(C.c_char_p)()
    n = C.c_int()
    _host_check(host_lib().qld_mdf_find_files(directory.encode(), hid.encode(),
                                              snips_per_file, total_snips, pos_wmin,
                                              pos_wmax, C.byref(files), C.byref(n)),
                f"chunks of {directory}")
    out = [files[i].decode() for i in range(n.value)]
    host_lib().qld_mdf_free_files(C.cast(files, C.c_void_p), n.value)
    return out


class MdfReader:
    """qld_mdf_reader: chunk files read and rows parsed once across calls."""

    def __init__(self, limit: int | None = None):
        self._r = host_lib().qld_mdf_reader_create()
        if limit is not None:
            host_lib().qld_mdf_reader_set_limit(self._r, limit)

    def table(self, files, pos_wmin, pos_wmax, words, pos_wset=1, nthreads=0) -> dict:
        return mdf_read_table(files, pos_wmin, pos_wmax, words, pos_wset, nthreads, reader=self)

    @property
    def cached_bytes(self) -> int:
        return int(host_lib().qld_mdf_reader_bytes(self._r))

    def close(self):
        if self._r:
            host_lib().qld_mdf_reader_free(self._r)
            self._r = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _table_dict(t: "_SnpTable") -> dict:
    r, w = t.rows, t.words
    snp = np.ctypeslib.as_array(C.cast(t.snp, C.POINTER(C.c_uint64)), (max(r, 1) * w,))
    pos = np.ctypeslib.as_array(C.cast(t.pos, C.POINTER(C.c_uint32)), (max(r, 1),))
    bc = np.ctypeslib.as_array(C.cast(t.bc, C.POINTER(C.c_uint32)), (max(r, 1),))
    idp = C.cast(t.id, C.POINTER(C.c_char_p))
    return {"snp": snp[:r * w].reshape(r, w).copy(), "pos": pos[:r].copy(),
            "bc": bc[:r].copy(), "id": [idp[i].decode() for i in range(r)]}


class VcfReader:
    """qld_vcf_reader over the parser's VCF chunks of one header file: rows
    as the correlator builds them without -mdf (writeToTable_x64), with the
    header's first and #CHROM lines and the sample selection of `sample_list`
    (None = every sample).  Reads draw unphased shuffles from the reader's own
    rand() stream, in call order."""

    def __init__(self, header_path: str, ploidy: int, sample_list: str | None = None,
                 seed: int | None = None):
        l1, ch = C.c_void_p(), C.c_void_p()
        rc = host_lib().qld_vcf_header_lines(header_path.encode(), C.byref(l1), C.byref(ch))
        if rc != QLD_OK:
            raise QldError(rc, f"header lines of {header_path}: "
                               f"{(host_lib().qld_vcf_last_error() or b'').decode()}")
        free = libc_free()
        self.line1, self.chrom_line = C.string_at(l1.value), C.string_at(ch.value)
        free(l1)
        free(ch)
        self._mask = C.c_void_p()
        ns, vc = C.c_int(), C.c_int()
        rc = host_lib().qld_vcf_sample_mask(self.chrom_line,
                                            sample_list.encode() if sample_list else None,
                                            ploidy, C.byref(self._mask), C.byref(ns),
                                            C.byref(vc))
        if rc != QLD_OK:
            raise QldError(rc, f"sample selection: "
                               f"{(host_lib().qld_vcf_last_error() or b'').decode()}")
        self.nsamples, self.valid_count, self.ploidy = ns.value, vc.value, int(ploidy)
        self._r = host_lib().qld_vcf_reader_create()
        if not self._r:
            raise QldError(QLD_ERR_NOMEM, "qld_vcf_reader_create")
        if seed is not None:
            host_lib().qld_vcf_reader_seed(self._r, int(seed))

    def table(self, files: list[str], pos_wmin: int, pos_wmax: int, snip_size: int,
              pos_wset: int = 1) -> dict:
        """Rows of `files` in [pos_wmin, pos_wmax] -> dict(snp, pos, bc, id)."""
        arr = (C.c_char_p * max(1, len(files)))(*[f.encode() for f in files])
        t = _SnpTable()
        rc = host_lib().qld_vcf_read_table(self._r, C.cast(arr, C.c_void_p), len(files),
                                           pos_wmin, pos_wmax, pos_wset, self.line1,
                                           self.chrom_line, self._mask, self.nsamples,
                                           self.valid_count, self.ploidy, int(snip_size),
                                           C.byref(t))
        if rc != QLD_OK:
            raise QldError(rc, f"reading VCF rows: "
                               f"{(host_lib().qld_vcf_last_error() or b'').decode()}")
        try:
            return _table_dict(t)
        finally:
            host_lib().qld_snp_table_free(C.byref(t))

    def close(self):
        if self._r:
            host_lib().qld_vcf_reader_free(self._r)
            self._r = None
        if self._mask:
            libc_free()(self._mask)
            self._mask = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def libc_free():
    f = C.CDLL(None).free
    f.argtypes = [C.c_void_p]
    f.restype = None
    return f


def mdf_read_table(files: list[str], pos_wmin: int, pos_wmax: int, words: int,
                   pos_wset: int = 1, nthreads: int = 0, reader: "MdfReader | None" = None) -> dict:
    """Rows of `files` in [pos_wmin, pos_wmax] -> dict(snp, pos, bc, id)."""
    arr = (C.c_char_p * max(1, len(files)))(*[f.encode() for f in files])
    t = _SnpTable()
    nt = nthreads or (os.cpu_count() or 1)
    if reader is None:
        rc = host_lib().qld_mdf_read_table(C.cast(arr, C.c_void_p), len(files), pos_wmin,
                                           pos_wmax, pos_wset, words, nt, C.byref(t))
    else:
        rc = host_lib().qld_mdf_reader_table(reader._r, C.cast(arr, C.c_void_p), len(files),
                                             pos_wmin, pos_wmax, pos_wset, words, nt,
                                             C.byref(t))
    _host_check(rc, "reading MDF rows")
    try:
        return _table_dict(t)
    finally:
        host_lib().qld_snp_table_free(C.byref(t))


def mdf_write(path: str, chrom: str, snp: np.ndarray, pos, ids, bc) -> None:
    snp = np.ascontiguousarray(snp, dtype=np.uint64)
    pos = np.ascontiguousarray(pos, dtype=np.uint32)
    bc = np.ascontiguousarray(bc, dtype=np.uint32)
    rows, words = snp.shape
    idarr = None
    if ids is not None:
        idarr = (C.c_char_p * max(1, rows))(*[i.encode() for i in ids])
    _host_check(host_lib().qld_mdf_write(path.encode(), chrom.encode(), _ptr(snp), _ptr(pos),
                                         None if idarr is None else C.cast(idarr, C.c_void_p),
                                         _ptr(bc), rows, words), f"writing {path}")


def format_float(val: float, precision: int = 9) -> str:
    buf = C.create_string_buffer(64)
    n = host_lib().qld_format_float(buf, C.c_float(val), precision)
    return buf.raw[:n].decode()


class DevBuffer:
    """Device memory owned by an Engine context."""

    def __init__(self, eng: "Engine", nbytes: int):
        self.eng = eng
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        eng._check(hip_lib().qld_dev_alloc(eng._ctx, self.nbytes, C.byref(p)))
        self.ptr = p.value or 0

    def upload(self, arr: np.ndarray) -> "DevBuffer":
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.eng._check(hip_lib().qld_h2d(self.eng._ctx, self.ptr, _ptr(arr), arr.nbytes))
        return self

    def download(self, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        self.eng._check(hip_lib().qld_d2h(self.eng._ctx, _ptr(out), self.ptr, out.nbytes))
        return out

    def download_at(self, offset: int, shape, dtype) -> np.ndarray:
        """Copy out `shape` elements starting `offset` bytes into the buffer."""
        out = np.empty(shape, dtype=dtype)
        assert offset >= 0 and offset + out.nbytes <= self.nbytes
        self.eng._check(hip_lib().qld_d2h(self.eng._ctx, _ptr(out), self.ptr + int(offset),
                                          out.nbytes))
        return out

    def fill(self, value_u32: int = 0):
        self.eng._check(hip_lib().qld_dev_fill(self.eng._ctx, self.ptr, value_u32,
                                               self.nbytes // 4))
        return self

    def free(self):
        if self.ptr:
            self.eng._check(hip_lib().qld_dev_free(self.eng._ctx, self.ptr))
            self.ptr = 0


class Engine:
    """One qld_ctx: a device, a stream, an engine choice and an r^2 mode."""

    def __init__(self, device=0, engine="auto", r2_mode="reference",
                 timing=False, hbm_budget=0):
        lib = hip_lib()
        cfg = Cfg()
        lib.qld_cfg_default(C.byref(cfg))
        cfg.device = device
        cfg.engine = ENGINES[engine] if isinstance(engine, str) else int(engine)
        cfg.r2_mode = R2_MODES[r2_mode] if isinstance(r2_mode, str) else int(r2_mode)
        cfg.timing = 1 if timing else 0
        cfg.hbm_budget = hbm_budget
        ctx = C.c_void_p()
        rc = lib.qld_gpu_init(C.byref(ctx), C.byref(cfg))
        if rc != QLD_OK:
            raise QldError(rc, lib.qld_gpu_last_error(None).decode())
        self._ctx = ctx
        self.cfg = cfg

    def close(self):
        if getattr(self, "_ctx", None):
            hip_lib().qld_gpu_release(self._ctx)
            self._ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != QLD_OK:
            raise QldError(rc, hip_lib().qld_gpu_last_error(self._ctx).decode())
        return rc

    @property
    def last_engine(self) -> str:
        return ENGINE_NAMES[hip_lib().qld_last_engine(self._ctx)]

    def correlate(self, A, bcA, B=None, bcB=None, snp_size=None,
                  triangular=None, results=True, counts=False):
        """Host-buffer all-pairs r^2 (qld_gpu_correlate)."""
        A = np.ascontiguousarray(A, dtype=np.uint64)
        bcA = np.ascontiguousarray(bcA, dtype=np.uint32)
        same = B is None
        if same:
            B, bcB = A, bcA
        B = np.ascontiguousarray(B, dtype=np.uint64)
        bcB = np.ascontiguousarray(bcB, dtype=np.uint32)
        m, W = A.shape
        n = B.shape[0]
        assert B.shape[1] == W
        if snp_size is None:
            snp_size = W * 64
        if triangular is None:
            triangular = same
        res = np.empty((n, m), dtype=np.float32) if results else None
        cnt = np.empty((n, m), dtype=np.uint32) if counts else None
        self._check(hip_lib().qld_gpu_correlate(
            self._ctx, _ptr(A), _ptr(bcA), m, _ptr(B), _ptr(bcB), n, W,
            int(snp_size), 1 if triangular else 0, _ptr(res), _ptr(cnt)))
        return res, cnt

    def correlate_hits(self, A, bcA, B=None, bcB=None, snp_size=None, upper=None,
                       r2limit=0.2, block_pairs=0):
        """Pairs the report prints, selected on the device (qld_correlate_hits):
        returns (offsets[n + 1] uint64, cols uint32, r2 float32) over all
        rows of B in order."""
        A = np.ascontiguousarray(A, dtype=np.uint64)
        bcA = np.ascontiguousarray(bcA, dtype=np.uint32)
        if B is None:
            B, bcB = A, bcA
            if upper is None:
                upper = True
        B = np.ascontiguousarray(B, dtype=np.uint64)
        bcB = np.ascontiguousarray(bcB, dtype=np.uint32)
        m, W = A.shape
        n = B.shape[0]
        if snp_size is None:
            snp_size = W * 64
        counts = np.zeros(n, dtype=np.uint64)
        parts = []

        def cb(_user, row0, rows, off, cols, r2):
            o = np.ctypeslib.as_array(off, (rows + 1,)).copy()
            h = int(o[-1])
            c = np.ctypeslib.as_array(cols, (h,)).copy() if h else np.zeros(0, np.uint32)
            v = np.ctypeslib.as_array(r2, (h,)).copy() if h else np.zeros(0, np.float32)
            counts[row0:row0 + rows] = np.diff(o)
            parts.append((row0, c, v))
            return 0

        fn = _HITS_FN(cb)
        self._check(hip_lib().qld_correlate_hits(
            self._ctx, _ptr(A), _ptr(bcA), m, _ptr(B), _ptr(bcB), n, W, int(snp_size),
            1 if upper else 0, C.c_float(r2limit), int(block_pairs),
            C.cast(fn, C.c_void_p), None))
        parts.sort(key=lambda t: t[0])
        offsets = np.zeros(n + 1, dtype=np.uint64)
        np.cumsum(counts, out=offsets[1:])
        cols = np.concatenate([p[1] for p in parts]) if parts else np.zeros(0, np.uint32)
        r2 = np.concatenate([p[2] for p in parts]) if parts else np.zeros(0, np.float32)
        return offsets, cols, r2

    def compact_dev(self, dres, ldr, rows, cols, row0, col0, upper, r2limit, dcounts,
                    doffsets, dcols, dr2, capacity):
        """qld_compact_dev on device buffers; returns the hit total."""
        def p(x):
            return None if x is None else (x.ptr if isinstance(x, DevBuffer) else int(x))
        total = C.c_uint64()
        self._check(hip_lib().qld_compact_dev(
            self._ctx, p(dres), int(ldr), rows, cols, row0, col0, 1 if upper else 0,
            C.c_float(r2limit), p(dcounts), p(doffsets), p(dcols), p(dr2), int(capacity),
            C.byref(total)))
        return total.value

    def alloc(self, nbytes: int) -> DevBuffer:
        return DevBuffer(self, nbytes)

    def correlate_dev(self, dA, dbcA, m, dB, dbcB, n, words, snp_size,
                      triangular, dres=None, ldr=0, dcnt=None, ldc=0):
        def p(x):
            return None if x is None else (x.ptr if isinstance(x, DevBuffer) else int(x))
        self._check(hip_lib().qld_correlate_dev(
            self._ctx, p(dA), p(dbcA), m, p(dB), p(dbcB), n, words, snp_size,
            1 if triangular else 0, p(dres), ldr or m, p(dcnt), ldc or m))

    def correlate_dev_noalloc(self, dA, dbcA, m, dB, dbcB, n, words, snp_size,
                              triangular, dres=None, ldr=0, dcnt=None, ldc=0):
        """qld_correlate_dev_noalloc: the same call without any allocation."""
        def p(x):
            return None if x is None else (x.ptr if isinstance(x, DevBuffer) else int(x))
        self._check(hip_lib().qld_correlate_dev_noalloc(
            self._ctx, p(dA), p(dbcA), m, p(dB), p(dbcB), n, words, snp_size,
            1 if triangular else 0, p(dres), ldr or m, p(dcnt), ldc or m))

    def sync(self):
        self._check(hip_lib().qld_sync(self._ctx))

    def timer_start(self):
        self._check(hip_lib().qld_timer_start(self._ctx))

    def timer_stop(self) -> float:
        ms = C.c_float()
        self._check(hip_lib().qld_timer_stop(self._ctx, C.byref(ms)))
        return ms.value

    def last_contract_ms(self) -> float:
        ms = C.c_float()
        self._check(hip_lib().qld_last_contract_ms(self._ctx, C.byref(ms)))
        return ms.value

    # ---- resident tables and window plans ---------------------------------
    def table(self, rows, bc, snp_size=None) -> "Table":
        return Table(self, rows, bc, snp_size)

    def plan_windows(self, table: "Table", windows: np.ndarray, out,
                     mode=BATCH_PERSISTENT) -> "Plan":
        """windows: WINDOW_DTYPE (triangles, qld_plan_windows) or
        PAIR_WINDOW_DTYPE (triangles and rectangles, qld_plan_pairs)."""
        return Plan(self, table, windows, out, mode)

    def set_fp4_operands(self, mode):
        """True / "packed" (1), False / "expanded" (0), or "direct" (2)."""
        m = {"expanded": 0, "packed": 1, "direct": 2}.get(mode, None) if isinstance(mode, str) \
            else (int(mode) if not isinstance(mode, bool) else (1 if mode else 0))
        self._check(hip_lib().qld_set_fp4_operands(self._ctx, m))

    def set_plain_tile(self, tile: int):
        self._check(hip_lib().qld_set_plain_tile(self._ctx, int(tile)))

    def set_timing(self, on: bool):
        """Per-call kernel events on (last_kernel_ms / last_contract_ms) or off."""
        self._check(hip_lib().qld_set_timing(self._ctx, 1 if on else 0))

    def set_ksplit(self, chunks: int):
        """Split-K chunks per 128-tile of plain calls: 0 by shape, 1 never, 2-64 forced."""
        self._check(hip_lib().qld_set_ksplit(self._ctx, int(chunks)))

    @property
    def last_ksplit(self) -> int:
        """Chunks per tile of the last plain MFMA call (1: not split)."""
        return hip_lib().qld_last_ksplit(self._ctx)

    def set_tile_cutoff(self, cutoff: int):
        self._check(hip_lib().qld_set_tile_cutoff(self._ctx, int(cutoff)))

    def set_plan_schedule(self, sched: str):
        """'waves': every wave claims wave blocks of its own (default);
        'tiles': a workgroup claims a whole tile."""
        self._check(hip_lib().qld_set_plan_schedule(self._ctx, {"tiles": 0, "waves": 1}[sched]))

    def set_plan_shared(self, shared: bool):
        """Whether persistent plans of single-region windows count each pair
        once (the QLD_BATCH_SHARED form); the autotuner sets it."""
        self._check(hip_lib().qld_set_plan_shared(self._ctx, 1 if shared else 0))

    @property
    def plan_shared(self) -> bool:
        return hip_lib().qld_get_plan_shared(self._ctx) == 1

    def set_plan_graphs(self, capture: bool):
        """Whether plans built afterwards are captured in a HIP graph."""
        self._check(hip_lib().qld_set_plan_graphs(self._ctx, 1 if capture else 0))

    @property
    def plan_schedule(self) -> str:
        return {0: "tiles", 1: "waves"}[hip_lib().qld_get_plan_schedule(self._ctx)]

    @property
    def tile_cutoff(self) -> int:
        return hip_lib().qld_get_tile_cutoff(self._ctx)

    def autotune_windows(self, table: "Table", windows: np.ndarray, out, reps=3):
        w = np.ascontiguousarray(windows, dtype=WINDOW_DTYPE)
        best = C.c_int()
        ms = C.c_float()
        optr = out.ptr if isinstance(out, DevBuffer) else int(out)
        self._check(hip_lib().qld_autotune_windows(self._ctx, table._t, _ptr(w), len(w),
                                                   optr, reps, C.byref(best), C.byref(ms)))
        return best.value, ms.value

    def last_kernel_ms(self) -> float:
        ms = C.c_float()
        self._check(hip_lib().qld_last_kernel_ms(self._ctx, C.byref(ms)))
        return ms.value


class Table:
    """A packed table resident in HBM, held in the layout the context's
    engine reads (qld_table_create / qld_table_create_dev)."""

    def __init__(self, eng: Engine, rows: np.ndarray, bc: np.ndarray, snp_size=None):
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        bc = np.ascontiguousarray(bc, dtype=np.uint32)
        self.eng = eng
        self.rows, self.words = rows.shape
        self.snp_size = int(snp_size if snp_size is not None else self.words * 64)
        t = C.c_void_p()
        eng._check(hip_lib().qld_table_create(eng._ctx, _ptr(rows), _ptr(bc), self.rows,
                                              self.words, self.snp_size, C.byref(t)))
        self._t = t

    @classmethod
    def from_device(cls, eng: Engine, rows_buf: "DevBuffer", bc_buf: "DevBuffer",
                    nrows: int, words: int, snp_size: int) -> "Table":
        """A table copied from packed rows and bit counts already in device
        buffers (qld_table_create_dev)."""
        self = cls.__new__(cls)
        self.eng = eng
        self.rows, self.words, self.snp_size = int(nrows), int(words), int(snp_size)
        t = C.c_void_p()
        eng._check(hip_lib().qld_table_create_dev(eng._ctx, rows_buf.ptr, bc_buf.ptr,
                                                  self.rows, self.words, self.snp_size,
                                                  C.byref(t)))
        self._t = t
        return self

    @property
    def nbytes(self) -> int:
        return hip_lib().qld_table_bytes(self._t)

    def free(self):
        if self._t:
            self.eng._check(hip_lib().qld_table_free(self.eng._ctx, self._t))
            self._t = None


class Plan:
    """A prepared window set (qld_plan_windows)."""

    def __init__(self, eng: Engine, table: Table, windows: np.ndarray, out, mode):
        pairs = windows.dtype == PAIR_WINDOW_DTYPE
        w = np.ascontiguousarray(windows, dtype=PAIR_WINDOW_DTYPE if pairs else WINDOW_DTYPE)
        self.eng = eng
        self.windows = w
        p = C.c_void_p()
        optr = out.ptr if isinstance(out, DevBuffer) else int(out)
        fn = hip_lib().qld_plan_pairs if pairs else hip_lib().qld_plan_windows
        eng._check(fn(eng._ctx, table._t, _ptr(w), len(w), optr, int(mode), C.byref(p)))
        self._p = p

    def run(self):
        self.eng._check(hip_lib().qld_plan_run(self.eng._ctx, self._p))

    def use_graph(self, flag: bool):
        self.eng._check(hip_lib().qld_plan_use_graph(self._p, 1 if flag else 0))

    def stats(self):
        t, p, n = C.c_int64(), C.c_int64(), C.c_int()
        self.eng._check(hip_lib().qld_plan_stats(self._p, C.byref(t), C.byref(p), C.byref(n)))
        return {"tiles": t.value, "pairs": p.value, "launches": n.value}

    def free(self):
        if self._p:
            self.eng._check(hip_lib().qld_plan_free(self.eng._ctx, self._p))
            self._p = None


def window_layout(starts, sizes) -> np.ndarray:
    """Window array with consecutive output blocks (size*size floats each)."""
    sizes = np.asarray(sizes, dtype=np.int64)
    w = np.zeros(len(sizes), dtype=WINDOW_DTYPE)
    w["start"] = starts
    w["size"] = sizes
    w["out_offset"] = np.concatenate([[0], np.cumsum(sizes * sizes)[:-1]]) if len(sizes) else []
    return w


class Server:
    """Continuous-batching request loop over resident tables (qld_server_*)."""

    def __init__(self, eng: Engine, out_budget=0, max_batch=0, streams=0):
        self.eng = eng
        cfg = ServerCfg(out_budget, max_batch, streams, 0)
        p = C.c_void_p()
        eng._check(hip_lib().qld_server_create(eng._ctx, C.byref(cfg), C.byref(p)))
        self._s = p

    def submit(self, table: Table, start: int, size: int, flags=REQ_KEEP, host_out=None):
        rid = C.c_int64()
        rc = hip_lib().qld_server_submit(self._s, table._t, int(start), int(size), int(flags),
                                         host_out, C.byref(rid))
        if rc != QLD_OK:
            raise QldError(rc, f"qld_server_submit({start}, {size}) failed")
        return rid.value

    def wait(self, rid: int, timeout_ms: float = -1.0) -> float:
        lat = C.c_double()
        rc = hip_lib().qld_server_wait(self._s, rid, timeout_ms, C.byref(lat))
        if rc != QLD_OK:
            raise QldError(rc, f"qld_server_wait({rid}) failed")
        return lat.value

    def fetch(self, rid: int, size: int) -> np.ndarray:
        out = np.empty((size, size), dtype=np.float32)
        rc = hip_lib().qld_server_fetch(self._s, rid, _ptr(out))
        if rc != QLD_OK:
            raise QldError(rc, f"qld_server_fetch({rid}) failed")
        return out

    def release(self, rid: int):
        rc = hip_lib().qld_server_release(self._s, rid)
        if rc != QLD_OK:
            raise QldError(rc, f"qld_server_release({rid}) failed")

    def stats(self):
        c, b, p = C.c_int64(), C.c_int64(), C.c_int64()
        rc = hip_lib().qld_server_stats(self._s, C.byref(c), C.byref(b), C.byref(p))
        return {"completed": c.value, "batches": b.value, "pairs": p.value, "status": rc}

    def close(self):
        if self._s:
            hip_lib().qld_server_destroy(self._s)
            self._s = None
