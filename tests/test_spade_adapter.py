"""The adapter of tests/cpp/spade_adapter.cpp (the code a SPADE user adds to
switch the hot path to spade_amd) compiles and links against the SPADE
headers and the library, and its device grid description and host layout
agree with SPADE's own objects (CPU only: the check needs no GPU).

Runs only where the SPADE sources are (they are not shipped with this
repository); built with g++ and the MPICH headers the SPADE include needs, as
oracle/ref/Makefile builds the golden-vector driver."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
MPI = "/opt/conda"
LIB = os.path.join(ROOT, "spade_amd", "libspade_amd.so")

pytestmark = pytest.mark.skipif(not (os.path.isdir(REF) and os.path.exists(os.path.join(MPI, "include", "mpi.h"))),
                                reason="SPADE sources or MPICH headers not present")


def test_adapter_compiles_links_and_matches_spade():
    if not os.path.exists(LIB):
        pytest.skip("libspade_amd.so is not built")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "spade_adapter")
        cmd = ["g++", "-std=c++20", "-O1", "-w", f"-I{REF}", f"-I{MPI}/include", f"-I{ROOT}/include",
               os.path.join(ROOT, "tests", "cpp", "spade_adapter.cpp"), "-o", exe,
               f"-L{ROOT}/spade_amd", "-lspade_amd", f"{MPI}/lib/libmpi.so",
               # the system C++ runtime first (the MPICH folder carries an older one)
               f"-Wl,-rpath,{ROOT}/spade_amd", "-Wl,-rpath,/usr/lib/x86_64-linux-gnu", f"-Wl,-rpath,{MPI}/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-3000:]
        run = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        print(run.stdout)
        assert run.returncode == 0, run.stdout + run.stderr
        assert "0 mismatches" in run.stdout
