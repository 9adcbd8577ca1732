"""Build recipe for the gfx950 library ``libshapes_hip.so``.

Every HIP translation unit is compiled with hipcc for gfx950 only; the host
generator code (``datagen.cpp``) with g++.  Objects go to
``shapeextraction_amd/_build/`` and the shared library lands next to this
file, where ``shapeextraction_amd._lib`` loads it.  ``build()`` always
recompiles.

Flags that matter for parity:
  -ffp-contract=off   no a*b+c -> fma contraction anywhere (the bit-exact
                      paths additionally use __fadd_rn/__fmul_rn/...)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions on the device
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libshapes_hip.so"
INCLUDE = PKG.parent / "include"

HIP_SOURCES = [
    "runtime.hip",
    "morton.hip",
    "sort.hip",
    "tree.hip",
    "octree.hip",
    "knn.hip",
    "normals.hip",
    "index.hip",
    "mlp.hip",
    "batch.hip",
]
CXX_SOURCES = ["datagen.cpp"]

ARCH = os.environ.get("SX_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    f"-I{INCLUDE}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall", f"-I{INCLUDE}"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    """Compile every source and link ``libshapes_hip.so``; returns its path."""
    BUILD.mkdir(exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    cmds = []
    objs = []
    for src in HIP_SOURCES:
        if not (CSRC / src).exists():
            continue
        obj = BUILD / (src + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    for src in CXX_SOURCES:
        obj = BUILD / (src + ".o")
        objs.append(obj)
        cmds.append(["g++", *CXX_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_run, c): c for c in cmds}
        for f in cf.as_completed(futs):
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lgomp"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def clean() -> None:
    if BUILD.exists():
        shutil.rmtree(BUILD)


if __name__ == "__main__":
    build()
