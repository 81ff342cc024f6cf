"""Host-only entry points of libqmx built with AddressSanitizer and
UndefinedBehaviorSanitizer (host code only; the device objects are the
regular build's) and driven by tests/sanitize/host_check.cpp: workspace
queries over every size, the forward-pass schedule with short and exact
capacities, ragged planning, MLP sizes, and the argument checks that return
before any HIP call.  Runs without a GPU.
"""

from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "qmprs_amd" / "csrc"
OBJ = ROOT / "qmprs_amd" / "_build"
DEVICE_OBJECTS = ["qmx_kernels.o", "qmx_linalg.o", "qmx_window.o", "qmx_mlp.o", "qmx_scratch.o",
                  "qmx_prep.o"]
HOST_SOURCES = [SRC / "qmx_api.cpp", SRC / "qmx_ragged.cpp", ROOT / "tests" / "sanitize" / "host_check.cpp"]
# each sanitizer flag applies to the host compilation only
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-sanitize-recover=all"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    return None


def test_host_entry_points_under_asan_ubsan(tmp_path):
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not available")
    missing = [o for o in DEVICE_OBJECTS if not (OBJ / o).exists()]
    if missing:
        pytest.skip(f"build libqmx first (missing {missing})")
    flags = ["-O1", "-g", "-fno-omit-frame-pointer", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument", f"-I{ROOT / 'include'}", f"-I{SRC}"] + SAN
    objs = []
    for src in HOST_SOURCES:
        obj = tmp_path / (src.stem + ".o")
        r = subprocess.run([hipcc, "-x", "hip", "-c", str(src), "-o", str(obj)] + flags,
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-4000:]
        objs.append(str(obj))
    exe = tmp_path / "host_check"
    r = subprocess.run([hipcc, "--offload-arch=gfx950"] + SAN[:4] + ["-o", str(exe)] + objs
                       + [str(OBJ / o) for o in DEVICE_OBJECTS], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    nm = subprocess.run(["nm", str(exe)], capture_output=True, text=True).stdout
    assert "__asan_report" in nm and "__ubsan_handle" in nm, "sanitizer runtime not linked"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "host checks passed" in r.stdout
