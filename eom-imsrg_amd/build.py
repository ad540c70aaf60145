"""Build recipe of the HIP library (gfx950) and of the Fortran C-ABI shim.

    python eom-imsrg_amd/build.py            # library + oracle
    python eom-imsrg_amd/build.py --no-oracle

Outputs (git-ignored, shipped in-tree):
    eom-imsrg_amd/eom_imsrg/libimsrg_hip.so
    oracle/_build/liboracle.so      (test infrastructure, built by oracle/Makefile)
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "eom_imsrg")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libimsrg_hip.so")

ARCH = "gfx950"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# packed FP32 ops are disabled for every HIP translation unit
DEV_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I", os.path.join(ROOT, "include"), "-I", CSRC]

SOURCES = ["basis.cpp", "interaction.cpp", "hf.cpp", "kernels.hip", "kernels_adj.hip", "gemm.hip", "tensor.hip",
           "capi.hip"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, *DEV_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
               "-x", "c++", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build_library(verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_oracle(verbose: bool = False) -> None:
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built oracle/_build/liboracle.so")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-oracle", action="store_true")
    a = ap.parse_args(argv)
    build_library(verbose=True)
    if not a.no_oracle:
        build_oracle(verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
