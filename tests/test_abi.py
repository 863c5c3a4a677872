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


def test_builtin_rk_tables_match_the_oracle():
    """spd_rk_table_builtin (host-only call) gives the tables the oracle's golden
    tests run (explicit.h:41-116)."""
    from spade_amd import _abi
    import spade_amd as sa
    from oracle import pyoracle as po
    sa.load()
    names = {"rk2": sa.rk2, "rk2hs": sa.rk2hs, "rk4": sa.rk4, "ssprk34": sa.ssprk34, "ssprk3": sa.ssprk3_table,
             "ssprk3hs": sa.ssprk3hs, "rk38r": sa.rk38r}
    from fractions import Fraction as F

    def frac(r):
        return F(0) if r.den == 0 else F(r.num, r.den)
    for name, which in names.items():
        t = sa.rk_table(which)
        a, b, c, hs = po.RK_TABLES[name]
        assert t.stages == len(b) and bool(t.high_storage) == hs, name
        for i in range(t.stages):
            for j in range(_abi.RK_MAX_STAGES):
                want = F(*a[i][j]) if j < len(a[i]) else F(0)
                assert frac(t.a[i][j]) == want, (name, i, j)
            assert frac(t.b[i]) == F(*b[i]) and frac(t.c[i]) == F(*c[i]), (name, i)
