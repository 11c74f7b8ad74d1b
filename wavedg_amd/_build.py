"""Builds libwavedg_hip.so in-tree with hipcc for gfx950.

Usage: ``python -m wavedg_amd._build`` (or ``__graft_entry__.build()``).
Objects go to wavedg_amd/_build/, the library to wavedg_amd/libwavedg_hip.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.environ.get("WDG_BUILD_DIR", os.path.join(HERE, "_build"))
LIB = os.environ.get("WDG_LIB_OUT", os.path.join(HERE, "libwavedg_hip.so"))
ORDERS = range(2, 10)  # n_colloc = P + 1

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("WDG_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed-FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
] + os.environ.get("WDG_EXTRA_HIPFLAGS", "").split()
CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")


def build(jobs: int | None = None, verbose: bool = False) -> str:
    """Compile every translation unit and link the shared library; returns its path."""
    if not os.path.exists(HIPCC) and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found; cannot build the HIP library")
    os.makedirs(OUT_DIR, exist_ok=True)
    steps: list[tuple[list[str], str]] = []
    for n in ORDERS:
        obj = os.path.join(OUT_DIR, f"wave2d_kernels_n{n}.o")
        steps.append(([HIPCC, *HIP_FLAGS, f"-DWDG_N={n}", "-c", os.path.join(CSRC, "wave2d_kernels_n.hip"), "-o", obj], obj))
    for src in ("wave2d_capi.hip", "batch.hip", "cnn.hip", "divf.hip", "exact.hip", "advect.hip", "flux_rtc.hip"):
        obj = os.path.join(OUT_DIR, src.replace(".hip", ".o"))
        steps.append(([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj], obj))
    for src in ("host_basis.cpp", "host_mesh.cpp"):
        obj = os.path.join(OUT_DIR, src.replace(".cpp", ".o"))
        steps.append(([CXX, *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj], obj))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, cmd) for cmd, _ in steps]
        for f in futs:
            f.result()
    objs = [o for _, o in steps]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-lhiprtc", "-o", tmp])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_dropin_driver(verbose: bool = False) -> str:
    """Builds tests/cpp/dropin_driver.cpp, a C++ consumer of the dg::hip
    drop-in classes (C ABI only, no HIP headers), next to the objects."""
    root = os.path.dirname(HERE)
    os.makedirs(OUT_DIR, exist_ok=True)
    exe = os.path.join(OUT_DIR, "dropin_driver")
    _run([CXX, "-std=c++17", "-O2", "-Wall", "-Wextra", f"-I{root}", os.path.join(root, "tests", "cpp", "dropin_driver.cpp"),
          f"-L{os.path.dirname(LIB)}", "-lwavedg_hip", "-Wl,-rpath,$ORIGIN/..", "-o", exe])
    if verbose:
        print(f"built {exe}")
    return exe


if __name__ == "__main__":
    build(verbose=True)
    build_dropin_driver(verbose=True)
