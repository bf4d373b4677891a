"""Builds the gfx950 shared library libconicxx_hip.so in-tree with hipcc.

Every source under csrc/ is compiled (as HIP, offload arch gfx950) to an
object under _build/ and linked into conicxx_amd/libconicxx_hip.so.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libconicxx_hip.so")

SOURCES = ["symbolic.cpp", "kkt_device.hip", "ipm_kernels.hip", "solver.hip", "batch.hip", "capi.cpp", "convert.cpp"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CONICXX_OFFLOAD_ARCH", "gfx950")
CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"),
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + CFLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines()
                     if "warning" in l and "packed-fp32-ops" not in l)
    return obj, warn


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for _, w in results:
            if w:
                print(w, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_facade_test()
    return LIB


FACADE_SRC = os.path.join(ROOT, "tests", "cpp", "facade_qp.cpp")
FACADE_BIN = os.path.join(ROOT, "tests", "cpp", "facade_qp")


def build_facade_test():
    """C++ program that drives include/conicxx_amd.hpp over the built library."""
    if not os.path.exists(FACADE_SRC):
        return None
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"), FACADE_SRC,
           "-L", HERE, "-lconicxx_hip", "-Wl,-rpath,$ORIGIN/../../conicxx_amd", "-o", FACADE_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"facade test build failed:\n{r.stdout}\n{r.stderr}")
    return FACADE_BIN


if __name__ == "__main__":
    print(build(verbose=True))
