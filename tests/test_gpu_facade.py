"""The C++ facade (include/spade_amd.hpp) drives the library as a SPADE-style
driver would: tests/cpp/facade_demo.cpp (built by build()) runs a TGV
flux_div and 3 SSP-RK3 steps on the GPU; its outputs are checked against the
CPU oracle (1e-12 relative max-norm per variable; the power law enters)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import golden_io as gi
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "spade_amd", "_build", "facade_demo")


def test_facade_demo_matches_oracle():
    assert os.path.exists(EXE), "facade_demo not built (run __graft_entry__.build())"
    d = tempfile.mkdtemp()
    r = subprocess.run([EXE, d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    dt = float(r.stdout.split()[1])
    shape = (8, 20, 20, 20, 5)
    prim0 = np.fromfile(os.path.join(d, "facade_prim0.bin")).reshape(shape)
    rhs = np.fromfile(os.path.join(d, "facade_rhs.bin")).reshape(shape)
    step3 = np.fromfile(os.path.join(d, "facade_step3.bin")).reshape(shape)
    gamma = 1.4
    R = 1.0 / 0.01 / gamma
    L = 2.0 * np.pi
    g = po.make_grid((2, 2, 2), (16, 16, 16), 2, (0, 0, 0), (L, L, L))
    fl = po.make_flux(po.make_term(po.TERM_KEEP, order=2),
                      po.make_term(po.TERM_VISC, law=po.LAW_POWER, mu=1 / 1600, T_ref=1.0, power=0.76, prandtl=0.72))
    gas = po.Gas(gamma, R)
    ref = po.flux_div(g, gas, fl, prim0)
    assert np.all(gi.rel_maxnorm(gi.interior(rhs, 2), gi.interior(ref, 2)) <= 1e-12)
    q = prim0.copy()
    po.ssprk3_step(g, gas, fl, po.BC(0, 0.0), dt, q, 3)
    err = gi.rel_maxnorm(gi.interior(step3, 2), gi.interior(q, 2))
    print("facade 3 steps rel err", err)
    assert np.all(err <= 1e-12)
    # rk_integrator_t with rk4_t, a caller residual, the transform-aware advance
    rk4 = np.fromfile(os.path.join(d, "facade_rk4_step2.bin")).reshape(shape)
    q = prim0.copy()
    po.rk_transform_step(g, gas, fl, None, po.rk_table("rk4"), dt, q, steps=2)
    err = gi.rel_maxnorm(gi.interior(rk4, 2), gi.interior(q, 2))
    print("facade rk4 2 steps rel err", err)
    assert np.all(err <= 1e-12)
    # io::binary_write wrote the host image; binary_read restored it; output_vtk wrote every block
    with open(os.path.join(d, "facade_step3.bin"), "rb") as a, \
            open(os.path.join(d, "facade_step3_checkpoint.bin"), "rb") as b, \
            open(os.path.join(d, "facade_step3_readback.bin"), "rb") as c:
        img = a.read()
        assert b.read() == img and c.read() == img
    assert os.path.exists(os.path.join(d, "vtk", "step3.pvts"))
    assert sorted(os.listdir(os.path.join(d, "vtk", "data_step3"))) == [f"b{lb:09d}.vts" for lb in range(8)]
