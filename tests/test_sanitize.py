"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU
only): tests/cpp/sanitize_driver.cpp builds meshes (structured, perturbed,
L-shaped, the reference's meshes), their face tables for every order, the
basis tables, and runs the CPU oracle's operators, RK4, WaveHoltz,
advection and Euler terms on small meshes.  A sanitizer report or a failed
check ends the driver with a nonzero status."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUT = os.path.join(ROOT, "wavedg_amd", "_build", "sanitize")
SAN = ["-O1", "-g", "-fopenmp", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
       "-fno-omit-frame-pointer"]


def _build() -> str:
    os.makedirs(OUT, exist_ok=True)
    obj = os.path.join(OUT, "wdg_oracle.o")
    exe = os.path.join(OUT, "sanitize_driver")
    subprocess.run(["gcc", "-std=c11", *SAN, "-c", os.path.join(ROOT, "oracle", "wdg_oracle.c"), "-o", obj],
                   check=True, capture_output=True, text=True)
    subprocess.run(["g++", "-std=c++17", *SAN, f"-I{ROOT}", os.path.join(ROOT, "tests", "cpp", "sanitize_driver.cpp"),
                    os.path.join(ROOT, "wavedg_amd", "csrc", "host_mesh.cpp"),
                    os.path.join(ROOT, "wavedg_amd", "csrc", "host_basis.cpp"), obj, "-lm", "-o", exe],
                   check=True, capture_output=True, text=True)
    return exe


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="needs gcc and g++")
def test_host_code_under_address_and_undefined_sanitizers():
    exe = _build()
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "meshes")], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "sanitize driver ok" in r.stdout
