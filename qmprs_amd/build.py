"""Build libqmx.so (HIP kernels + C-ABI) for gfx950 in-tree with hipcc.

Usage: ``python -m qmprs_amd.build`` (also called by ``__graft_entry__.build``).
Objects are compiled in parallel into ``qmprs_amd/_build/`` and linked into
``qmprs_amd/libqmx.so``.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SRC = PKG / "csrc"
OUT = PKG / "libqmx.so"
SOURCES = ["qmx_kernels.hip", "qmx_linalg.hip", "qmx_window.hip", "qmx_mlp.hip", "qmx_scratch.hip", "qmx_prep.hip",
           "qmx_api.cpp",
           "qmx_ragged.cpp"]
ARCH = os.environ.get("QMX_OFFLOAD_ARCH", "gfx950")


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; libqmx needs the ROCm toolchain")


def _flags() -> list[str]:
    return [
        "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
        # keep the compiler from emitting packed FP32 instructions
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        f"-I{ROOT / 'include'}", f"-I{SRC}",
    ]


def build(verbose: bool = False) -> Path:
    hipcc = _hipcc()
    bdir = PKG / "_build"
    bdir.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = bdir / (Path(src).stem + ".o")
        objs.append(obj)
        cmds.append([hipcc, "-x", "hip", "-c", str(SRC / src), "-o", str(obj)] + _flags())

    def run(cmd):
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        if verbose and (r.stderr.strip()):
            print(r.stderr, file=sys.stderr)

    with cf.ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1)) as ex:
        list(ex.map(run, cmds))
    tmp = OUT.with_suffix(".so.tmp")
    link = [hipcc, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp)] + [str(o) for o in objs]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(link)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
