"""Builds libnbfmm_amd.so (HIP for gfx950) in-tree with hipcc.

    python -m nbfmm_amd.build          # or nbfmm_amd.build.build()

Each csrc/*.hip translation unit is compiled to an object in
nbfmm_amd/_build/ and linked into nbfmm_amd/_build/libnbfmm_amd.so; the
command-line program csrc/nbfmm_main.cpp is linked against it as
nbfmm_amd/_build/nbfmm_main.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libnbfmm_amd.so")
MAIN = os.path.join(OUT, "nbfmm_main")
ROOT = os.path.dirname(HERE)

ARCH = os.environ.get("NBFMM_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled: hipcc would otherwise form
# v_pk_{add,mul,fma}_f32 from ordinary float code.
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]

SOURCES = ["sort.hip", "tree.hip", "p2p.hip", "adjoint.hip", "prune.hip", "model.hip", "display.hip", "capi.hip",
           "noalloc.hip"]


def _compile(src: str) -> str:
    obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), 8)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    tmp = MAIN + ".tmp"
    cmd = [HIPCC, *FLAGS, os.path.join(CSRC, "nbfmm_main.cpp"), "-o", tmp, "-L", OUT, "-lnbfmm_amd",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"nbfmm_main failed to build:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, MAIN)
    if verbose:
        print(f"built {LIB} and {MAIN}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
