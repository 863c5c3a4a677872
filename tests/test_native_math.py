"""Host checks of the FP64 kernel's numerical building blocks (CPU only).

Compiles tests/native/math_check.cpp with g++ (quad precision from libquadmath
as the exact reference) and checks:
* the binomial-series x^n used for the power-law viscosity stays within one
  ulp of x^n and is correctly rounded in the large majority of cases, with the
  full range |x - 1| <= 1/16 for the usual exponents;
* division by a constant with Markstein's correction step equals IEEE division.
"""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def result():
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tempfile.mkdtemp()
    exe = os.path.join(d, "math_check")
    src = os.path.join(HERE, "native", "math_check.cpp")
    r = subprocess.run(["g++", "-O2", "-std=c++17", src, "-o", exe, "-lquadmath"], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("libquadmath not available: " + r.stderr[-200:])
    out = subprocess.run([exe], capture_output=True, text=True, check=True, timeout=120).stdout.split()
    shutil.rmtree(d, ignore_errors=True)
    return [float(x) for x in out]


def test_series_pow_within_one_ulp(result):
    pn, pbad, maxulp, full_range = result[0], result[1], result[2], result[5]
    assert pn > 1100000
    assert pbad <= 0.05 * pn, pbad
    assert maxulp < 1.0, maxulp
    assert full_range == 6


def test_markstein_division_is_ieee(result):
    dn, dbad = result[3], result[4]
    assert dn == 1500000
    assert dbad == 0
