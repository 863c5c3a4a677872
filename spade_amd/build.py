"""Builds libspade_amd.so in-tree with hipcc for gfx950.

Every HIP translation unit is compiled with
  * -ffp-contract=off: floating-point contraction is enabled only where a
    kernel asks for it with a pragma, so the FP64 parity path keeps the
    reference's rounding sequence;
  * -Xclang -target-feature -Xclang -packed-fp32-ops: no packed FP32 VALU
    instructions are generated.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libspade_amd.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = ["api.cpp", "io.cpp", "rk_host.cpp", "flux_generic.hip", "flux_fast.hip", "grid_ops.hip", "rk.hip",
           "batch.hip", "batch_host.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-result"]


def _compile(src, obj, verbose):
    cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True, jobs=None):
    """Compile every source and link libspade_amd.so (always rebuilds)."""
    os.makedirs(OBJDIR, exist_ok=True)
    objs = [os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    procs = []
    for src, obj in zip(SOURCES, objs):
        cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
    failed = [src for src, p in procs if p.wait() != 0]
    if failed:
        raise RuntimeError("hipcc failed for: " + ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_facade_demo(verbose)
    return LIB


def build_facade_demo(verbose=True):
    """The C++ facade example (tests/cpp/facade_demo.cpp), linked against the library."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "facade_demo.cpp")
    exe = os.path.join(OBJDIR, "facade_demo")
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-I", os.path.join(root, "include"), src,
           "-o", exe, "-L", HERE, "-lspade_amd", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return exe


def clean():
    shutil.rmtree(OBJDIR, ignore_errors=True)
    if os.path.exists(LIB):
        os.remove(LIB)


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
