"""Resource properties of every kernel in libspade_amd.so as built here for
gfx950 (CPU test: reads the library's own code objects, no GPU needed).

* no kernel has a dynamic stack, and no kernel calls a function: every term
  of a flux is inlined into the kernel that evaluates it (compile-time
  specialised kernels, and the run-time dispatch of the generic kernel with
  its term functions forced inline), so each kernel's private segment has the
  fixed size the code object declares;
* the per-lane private (scratch) segment stays small (register spills and the
  few stencil arrays indexed at run time).
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "spade_amd", "libspade_amd.so")
LLVM = "/opt/rocm/llvm/bin"
MAX_PRIVATE_BYTES = 512


def _code_objects(tmp_path):
    if not os.path.exists(LIB):
        pytest.skip("libspade_amd.so is not built")
    if not os.path.exists(os.path.join(LLVM, "llvm-objdump")):
        pytest.skip("llvm-objdump not available")
    shutil.copy(LIB, tmp_path / "lib.so")
    subprocess.run([os.path.join(LLVM, "llvm-objdump"), "--offloading", str(tmp_path / "lib.so")], check=True,
                   capture_output=True, cwd=tmp_path)
    objs = sorted(p for p in tmp_path.iterdir() if p.name.endswith("gfx950"))
    assert objs, "no gfx950 code object in the library"
    return objs


def _kernels(obj):
    """(name, uses_dynamic_stack, private_segment_fixed_size) of every kernel."""
    out = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes", str(obj)], check=True,
                         capture_output=True, text=True).stdout
    rows, cur = [], None
    for line in out.splitlines():
        m = re.search(r"\.name:\s+(\S+)", line)
        if m and not m.group(1).endswith(".kd"):
            cur = {"name": m.group(1)}
            rows.append(cur)
            continue
        m = re.search(r"\.(uses_dynamic_stack|private_segment_fixed_size):\s+(\S+)", line)
        if m and cur is not None:
            cur[m.group(1)] = m.group(2)
    return rows


def test_no_dynamic_stack_no_calls_small_scratch(tmp_path):
    objs = _code_objects(tmp_path)
    names = []
    for obj in objs:
        for k in _kernels(obj):
            names.append(k["name"])
            assert k.get("uses_dynamic_stack") == "false", k
            assert int(k.get("private_segment_fixed_size", "0")) <= MAX_PRIVATE_BYTES, k
        dis = subprocess.run([os.path.join(LLVM, "llvm-objdump"), "-d", "--no-show-raw-insn", str(obj)],
                             check=True, capture_output=True, text=True).stdout
        calls = [l for l in dis.splitlines() if "s_swappc" in l]
        assert not calls, (obj.name, calls[:5])
    # the kernels of the flux path are all there
    for want in ("k_flux_fast", "k_flux_div_generic", "k_flux_div_spec", "k_rk_transform", "k_exchange"):
        assert any(want in n for n in names), want
