import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def golden_cases():
    return sorted(glob.glob(os.path.join(GOLDEN, "case_*.npz")))


def load_case(path):
    with np.load(path, allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    d["report"] = d["report"].tobytes().decode()
    d["name"] = os.path.basename(path)[5:-4]
    return d


def unpack_bits(table: np.ndarray, n_samples: int) -> np.ndarray:
    """Packed uint64 rows (MSB-first) -> 0/1 int64 matrix rows x n_samples."""
    t = np.ascontiguousarray(table, dtype=np.uint64)
    be = t.astype(">u8").view(np.uint8).reshape(t.shape[0], -1)
    return np.unpackbits(be, axis=1)[:, :n_samples].astype(np.int64)


def pack_bits(bits: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """0/1 matrix rows x N -> (packed uint64 rows MSB-first, bit counts)."""
    bits = np.asarray(bits, dtype=np.uint8)
    rows, N = bits.shape
    W = (N + 63) // 64
    padded = np.zeros((rows, W * 64), dtype=np.uint8)
    padded[:, :N] = bits
    be = np.packbits(padded, axis=1).view(">u8").astype(np.uint64)
    return np.ascontiguousarray(be), bits.sum(axis=1).astype(np.uint32)


@pytest.fixture(scope="session")
def engine_factory():
    import qld_amd
    made = []

    def make(**kw):
        # fp4 operand sources: "mfma_fp4" = the default (packed rows straight
        # into registers), "mfma_fp4d" = the same, named; "mfma_fp4p" = packed
        # rows expanded into LDS; "mfma_fp4x" = the expanded operand copy
        # streamed into LDS.
        mode = {"mfma_fp4p": "packed", "mfma_fp4d": "direct",
                "mfma_fp4x": "expanded"}.get(kw.get("engine"))
        if mode:
            kw["engine"] = "mfma_fp4"
        e = qld_amd.Engine(**kw)
        if mode:
            e.set_fp4_operands(mode)
        made.append(e)
        return e
    yield make
    for e in made:
        e.close()
