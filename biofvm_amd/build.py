"""Builds libbiofvm_amd.so (hand-written gfx950 HIP kernels + the C ABI),
libbiofvm_facade.so (the BioFVM-compatible C++ host façade over it) and
libbfvm_ceiling.so (the bandwidth probes of `bench.py --ceiling`) in-tree.

    python -m biofvm_amd.build          # or biofvm_amd.build.build()

The library is compiled with hipcc for gfx950 only.  Packed FP32 ALU ops are
disabled for the device code (``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""
import concurrent.futures
import os
import shlex
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libbiofvm_amd.so")

SOURCES = ["capi.hip", "sources.hip", "lod.hip", "ensemble.hip", "adjoint.hip", "io.hip", "agents.hip", "reduce.hip",
           "sort.hip"]
HEADERS = ["bfvm_internal.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def build(verbose=False):
    """Compiles every source and links the shared library; raises on failure."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(shlex.quote(c) for c in cmd), flush=True)
        cmds.append(cmd)
        objs.append(obj)
    # one compiler process per source, at most as many as the host has cores
    workers = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0")) or os.cpu_count() or 1, 16))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for proc in pool.map(lambda c: subprocess.run(c), cmds):
            if proc.returncode != 0:
                raise subprocess.CalledProcessError(proc.returncode, proc.args)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_facade(verbose)
    build_ceiling(verbose)
    return LIB


CEILING_LIB = os.path.join(HERE, "libbfvm_ceiling.so")


def build_ceiling(verbose=False):
    """The bandwidth-ceiling kernels `bench.py --ceiling` times (a library of
    their own: nothing in the product library calls them)."""
    tmp = CEILING_LIB + ".tmp"
    cmd = [HIPCC, *CFLAGS, "-shared", os.path.join(CSRC, "ceiling.hip"), "-o", tmp]
    if verbose:
        print(" ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, CEILING_LIB)
    return CEILING_LIB


FACADE_DIR = os.path.join(HERE, "facade")
FACADE_LIB = os.path.join(HERE, "libbiofvm_facade.so")


def build_facade(verbose=False):
    """The BioFVM-compatible C++ host façade (plain g++, links the C ABI library)."""
    tmp = FACADE_LIB + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall",
           f"-I{INCLUDE}", f"-I{FACADE_DIR}", os.path.join(FACADE_DIR, "BioFVM_gpu.cpp"),
           os.path.join(FACADE_DIR, "BioFVM_MultiCellDS_gpu.cpp"),
           os.path.join(FACADE_DIR, "BioFVM_vector_utilities.cpp"),
           os.path.join(FACADE_DIR, "BioFVM_matlab_io.cpp"),
           f"-L{HERE}", "-lbiofvm_amd", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    if verbose:
        print(" ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, FACADE_LIB)
    return FACADE_LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
