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

HIP_SOURCES = ["lagrange.hip", "sweep_wave.hip", "sweep_walk.hip", "sweep_cell.hip", "remap.hip", "eucclhyd.hip",
               "context.hip"]
HOST_SOURCES = ["host/donnees.cc", "host/json_input.cc", "host/init.cc", "host/init_euc.cc", "host/output.cc"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]
# per-source device flags: the column walker loops over rows, and
# machine-level hoisting of constants out of that loop would hold registers
# through every row
SOURCE_FLAGS = {"sweep_walk.hip": ["-mllvm", "-disable-machine-licm"]}


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
        jobs.append((src, obj, [cc, "-c", src, "-o", obj] + COMMON + DEVICE + SOURCE_FLAGS.get(s, [])))
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
    if os.path.exists(REFERENCE_MAIN):
        build_reference_main()
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


# The reference's own main (main.cc), compiled against include/mahyco.hpp to
# show that its caller needs no change beyond include lines and namespace
# names.  Built only where the reference sources are present; the source is
# piped to the compiler, never written into this tree.
REFERENCE_MAIN = "/root/reference/main.cc"
REFMAIN_EXE = os.path.join(BIN_DIR, "mahyco_refmain")
_STD_HEADERS = ("<cstdlib>", "<iostream>", "<string>")


def reference_main_source(text):
    """The reference's main.cc with its include lines replaced by this
    library's header (standard headers kept), `using namespace nablalib`
    by `using namespace mahyco`, and the Kokkos:: qualifier of the runtime
    lifecycle calls by mahyco::.  Returns (source, [(line no, old, new)])."""
    out, changed, header_done = [], [], False
    for n, line in enumerate(text.splitlines(), 1):
        new = line
        if line.lstrip().startswith("#include"):
            if any(h in line for h in _STD_HEADERS):
                new = line
            elif not header_done:
                new, header_done = '#include "mahyco.hpp"', True
            else:
                new = ""
        elif line.strip() == "using namespace nablalib;":
            new = "using namespace mahyco;"
        elif "Kokkos::" in line:
            new = line.replace("Kokkos::", "mahyco::")
        if new != line:
            changed.append((n, line, new))
        out.append(new)
    return "\n".join(out) + "\n", changed


def build_reference_main(exe=None, main_path=REFERENCE_MAIN):
    """Compiles the reference's main.cc (rewritten by reference_main_source)
    against include/mahyco.hpp and links it to the in-tree library."""
    exe = exe or REFMAIN_EXE
    with open(main_path) as fh:
        src, _ = reference_main_source(fh.read())
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    cxx = shutil.which("g++") or shutil.which("c++")
    tmp = exe + ".tmp"
    # in-tree: found relative to the executable (the tree moves as a whole);
    # elsewhere (a test's scratch directory): the library's absolute path
    in_tree = os.path.abspath(os.path.dirname(exe)) == os.path.abspath(BIN_DIR)
    rpath = "$ORIGIN/../lib" if in_tree else os.path.abspath(LIB_DIR)
    cmd = [cxx, "-std=c++17", "-O2", "-I", INCLUDE, "-x", "c++", "-", "-o", tmp, "-L", LIB_DIR, "-lmahyco_hip",
           "-Wl,-rpath," + rpath]
    r = subprocess.run(cmd, input=src, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("reference main build failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, exe)
    return exe


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
