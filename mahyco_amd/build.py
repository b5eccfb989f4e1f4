"""Builds libmahyco_hip.so (HIP kernels for gfx950 + host C++ + C ABI) in-tree.

The library lands in mahyco_amd/lib/; every call recompiles.  Device code is
compiled with -ffp-contract=off so FP64 results follow the reference's
operation order bit for bit, and with packed FP32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libmahyco_hip.so")
BIN_DIR = os.path.join(HERE, "bin")
EXE = os.path.join(BIN_DIR, "mahyco")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
ARCH = os.environ.get("MAHYCO_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["lagrange.hip", "sweep_wave.hip", "sweep_walk.hip", "sweep_cell.hip", "remap.hip", "context.hip"]
HOST_SOURCES = ["host/donnees.cc", "host/init.cc", "host/output.cc"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(args):
    src, obj, cmd = args
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compiling {src} failed:\n{r.stdout}\n{r.stderr}")
    return r.stderr


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    cc = hipcc()
    jobs = []
    for s in HIP_SOURCES:
        src = os.path.join(SRC, s)
        obj = os.path.join(OBJ, s.replace("/", "_") + ".o")
        jobs.append((src, obj, [cc, "-c", src, "-o", obj] + COMMON + DEVICE))
    for s in HOST_SOURCES:
        src = os.path.join(SRC, s)
        obj = os.path.join(OBJ, s.replace("/", "_") + ".o")
        jobs.append((src, obj, [cc, "-c", src, "-o", obj] + COMMON))
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        logs = list(ex.map(_compile, jobs))
    if verbose:
        for l in logs:
            if l.strip():
                print(l)
    tmp = LIB + ".tmp"
    cmd = [cc, "-shared", "-o", tmp] + [j[1] for j in jobs] + [f"--offload-arch={ARCH}", "-fPIC"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    build_driver()
    return LIB


def build_driver():
    """The mahyco executable (C++ drivers of include/mahyco.hpp over the C
    ABI), linked against the in-tree library."""
    os.makedirs(BIN_DIR, exist_ok=True)
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        raise RuntimeError("no host C++ compiler for the driver executable")
    tmp = EXE + ".tmp"
    cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-I", INCLUDE, os.path.join(SRC, "drivers", "main.cc"), "-o", tmp,
           "-L", LIB_DIR, "-lmahyco_hip", "-Wl,-rpath,$ORIGIN/../lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, EXE)
    return EXE


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
