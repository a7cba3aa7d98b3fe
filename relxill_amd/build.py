"""Builds librelxill_gpu.so (HIP for gfx950) in-tree, plus the oracle library.

Usage: python -m relxill_amd.build   (or relxill_amd.build.build_all())
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "librelxill_gpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# no packed FP32 instructions (v_pk_*_f32) anywhere in the device code
NO_PK = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CFLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC] + NO_PK

SOURCES = ["tables.cpp", "fits.cpp", "engine.cpp", "capi.cpp", "server.cpp", "kernels.hip", "noalloc.hip", "nthcomp.hip"]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    cmd = [HIPCC, "-c", path, "-o", obj] + CFLAGS
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "hip", "-c", path, "-o", obj] + CFLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build_library(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-o", tmp] + objs + [f"--offload-arch={ARCH}", "-lpthread", "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_oracle(verbose=False):
    d = os.path.join(ROOT, "oracle")
    r = subprocess.run(["make", "-C", d, "-B", "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", os.path.join(d, "liboracle.so"))


def build_all(verbose=False):
    build_oracle(verbose)
    build_library(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
