"""Build libgpmood.so (HIP kernels for gfx950 + C ABI) in-tree.

hipcc compiles the kernels and the C ABI; the quad-precision GP precompute is
host-only C++ (it uses __float128 and libquadmath) and is compiled with g++.
Every HIP translation unit is built with -ffp-contract=off (no FMA
contraction, so the FP64 arithmetic rounds as in the reference) and without
packed FP32 instructions.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgpmood.so")
ARCH = os.environ.get("GPM_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip")) + ["gpm_capi.cpp", "gpm_dispatch.cpp"]
HOST_SOURCES = ["gp_tables.cpp"]
IO_LIB = os.path.join(HERE, "libgpmood_io.so")
# HDF5 C library of the reference's file formats (libgpmood_io.so): HDF5_DIR,
# else the image's conda prefix
HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-3] if "-o" in cmd else cmd[-1]))


def build(verbose=True, jobs=None, out=None, defines=(), combos=None):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    for src in HOST_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        cmds.append(["g++", "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    dflags = [f"-D{d}" for d in defines]
    tag = "_".join(d.replace("=", "") for d in defines)
    for src in HIP_SOURCES:
        if combos is not None and src.startswith("gpm_k_m") and src not in combos:
            continue
        obj = os.path.join(objdir, src + (("." + tag) if tag else "") + ".o")
        cmds.append([HIPCC] + HIP_FLAGS + dflags + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(run, cmds))
    lib = out or LIB
    tmp = lib + ".tmp"
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lquadmath"])
    os.replace(tmp, lib)
    return lib


def build_io(verbose=True):
    """libgpmood_io.so: the reference's HDF5 formats (host code).  Skipped,
    with a message, when the HDF5 C headers are absent."""
    inc, lib = os.path.join(HDF5_DIR, "include"), os.path.join(HDF5_DIR, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        if verbose:
            print(f"build: {inc}/hdf5.h not found; libgpmood_io.so (HDF5 formats) not built", file=sys.stderr)
        return None
    tmp = IO_LIB + ".tmp"
    run(["g++", "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-shared", "-I" + inc,
         os.path.join(CSRC, "gpm_io.cpp"), "-o", tmp, "-L" + lib, "-lhdf5", "-Wl,-rpath," + lib])
    os.replace(tmp, IO_LIB)
    return IO_LIB


if __name__ == "__main__":
    build()
    build_io()
    sys.exit(0)
