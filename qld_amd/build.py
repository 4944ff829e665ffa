"""Build the native libraries in-tree.

libqld_hip.so   HIP kernels + C ABI (hipcc, gfx950 only)
libqld_host.so  C host side: synthetic data, report writer, drop-in shim

Run as ``python -m qld_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIBDIR = os.path.join(PKG, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # Packed-FP32 instructions are kept out of the device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # The r^2 epilogue must round exactly where the source says.
    "-ffp-contract=off",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-Wno-unused-value",
]
HOST_FLAGS = ["-O3", "-march=x86-64-v3", "-fPIC", "-shared", "-fopenmp",
              "-ffp-contract=off", "-std=gnu11", "-Wall", "-Wextra",
              "-Wno-unused-parameter"]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(PKG, "csrc", "*.hip")))
    out = os.path.join(LIBDIR, "libqld_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, *srcs])
    os.replace(tmp, out)
    return out


def build_host() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(PKG, "host", "*.c")))
    out = os.path.join(LIBDIR, "libqld_host.so")
    tmp = out + ".tmp"
    _run(["gcc", *HOST_FLAGS, f"-I{os.path.join(ROOT, 'include')}", "-o", tmp,
          *srcs, f"-L{LIBDIR}", "-lqld_hip", f"-Wl,-rpath,$ORIGIN", "-lz", "-lm"])
    os.replace(tmp, out)
    return out


def build_cli() -> str:
    """qLD-compute: the drop-in command line over libqld_host + libqld_hip."""
    out = os.path.join(LIBDIR, "qLD-compute")
    tmp = out + ".tmp"
    _run(["gcc", "-O2", "-std=gnu11", "-fopenmp", "-Wall", "-Wextra", "-Wno-unused-parameter",
          f"-I{os.path.join(ROOT, 'include')}", "-o", tmp,
          os.path.join(PKG, "cli", "qld_compute.c"), f"-L{LIBDIR}", "-lqld_host",
          "-lqld_hip", "-Wl,-rpath,$ORIGIN", "-lz", "-lm"])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_hip()
    build_host()
    build_cli()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "host": build_host, "cli": build_cli}[w]()
