"""The C ABI library loads and exports every entry point include/*.h declares.

CPU only: no compute call is made (there may be no GPU here)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(spd_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    assert len(_declared()) >= 20


@pytest.mark.parametrize("name", _declared())
def test_symbol_exported(name):
    from spade_amd import _abi
    lib = ctypes.CDLL(_abi.LIB_PATH)
    assert hasattr(lib, name), name


def test_python_signatures_cover_header():
    from spade_amd import _abi
    assert set(_declared()) <= set(_abi.SIGNATURES)
