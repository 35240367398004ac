"""Build the gfx950 kernel library libcsrst.so in-tree with hipcc.

The library is one shared object with the C ABI declared in
include/csrst_hip.h.  The two-grid step (csrst_ref.hip) is compiled without
FMA contraction (-ffp-contract=off) so that it is bit-identical to the
reference arithmetic; the other sources contract to FMAs.  Packed-FP32
instruction selection is disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcsrst.so")
SOURCES = ["csrst_ref.hip", "csrst_mw.hip", "csrst_poisson.hip", "csrst_unet.hip", "csrst_sf.hip"]
# The two-grid step and the stateless ShapeFunction / ScatterGatherImpl
# kernels must stay bit-identical to the reference arithmetic, so their
# translation units forbid FMA contraction.  The production window, the
# multigrid and the network are tolerance-checked and use fused multiply-adds,
# except inside `#pragma clang fp contract(off)` scopes (the reference's float
# fast inverse square root), which -ffp-contract=fast-honor-pragmas keeps.
CONTRACT = {"csrst_ref.hip": "off", "csrst_sf.hip": "off"}
ARCH = os.environ.get("CSRST_ARCH", "gfx950")


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c
    raise RuntimeError("hipcc not found")


def flags(contract="off"):
    return ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=" + contract,
            "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-I" + CSRC, "-I" + os.path.join(HERE, "..", "include")]


def build(verbose=True, jobs=4):
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = []
    procs = []
    for s in srcs:
        obj = os.path.join(CSRC, s.replace(".hip", ".o"))
        cmd = [hipcc()] + flags(CONTRACT.get(s, "fast-honor-pragmas")) + ["-c", os.path.join(CSRC, s), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), s))
        objs.append(obj)
        if len(procs) >= jobs:
            p, name = procs.pop(0)
            if p.wait() != 0:
                raise RuntimeError("compile failed: " + name)
    for p, name in procs:
        if p.wait() != 0:
            raise RuntimeError("compile failed: " + name)
    tmp = LIB + ".tmp%d" % os.getpid()
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    build_driver(verbose)
    build_h5part(verbose)
    return LIB


def build_driver(verbose=True):
    """The C++ front end's command-line driver (host/csrst_main.cpp), linked
    against libcsrst.so in this directory."""
    exe = os.path.join(HERE, "csrst")
    tmp = exe + ".tmp%d" % os.getpid()
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I" + os.path.join(HERE, "..", "include"),
           os.path.join(HERE, "host", "csrst_main.cpp"), "-o", tmp, "-L" + HERE, "-lcsrst",
           "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, exe)
    return exe


HDF5_ROOT = "/opt/conda"


def build_h5part(verbose=True):
    """The H5Part converter (host/csrst_h5part.c), plain C over the HDF5 C
    library; skipped when HDF5's headers are not installed."""
    if not os.path.exists(os.path.join(HDF5_ROOT, "include", "hdf5.h")):
        if verbose:
            print("HDF5 not found under %s: csrst_h5part not built" % HDF5_ROOT, flush=True)
        return None
    exe = os.path.join(HERE, "csrst_h5part")
    tmp = exe + ".tmp%d" % os.getpid()
    lib = os.path.join(HDF5_ROOT, "lib")
    cmd = ["gcc", "-std=gnu99", "-O2", "-Wall", "-Wextra", "-I" + os.path.join(HDF5_ROOT, "include"),
           os.path.join(HERE, "host", "csrst_h5part.c"), "-o", tmp, "-L" + lib, "-lhdf5", "-Wl,-rpath," + lib]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, exe)
    return exe


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
