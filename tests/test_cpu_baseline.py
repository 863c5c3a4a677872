"""The CPU baseline bench.py reports beside the GPU figure (CPU only): the
oracle's restatement of flux_div_basic timed on one core and on several
cores with OpenMP over blocks.  The threaded loop writes each block's cells
from one thread only, so its result is the serial result bit for bit.
"""
import numpy as np

import bench
from oracle import pyoracle as po


def _state(b=2):
    L, p0 = 2.0 * np.pi, 100.0 / 1.4
    nb, nc, ng = (b,) * 3, (16, 16, 16), 2
    g = po.make_grid(nb, nc, ng, (0, 0, 0), (L, L, L))
    q = np.zeros(po.shape_of(g))
    q[:, ng:-ng, ng:-ng, ng:-ng] = bench.tgv_interior(nb, nc, L, p0)
    po.exchange(g, q)
    fl = po.make_flux(po.make_term(po.TERM_KEEP, order=2),
                      po.make_term(po.TERM_VISC, law=po.LAW_POWER, mu=1 / 1600, T_ref=1.0, power=0.76, prandtl=0.72))
    return g, q, fl, po.Gas(1.4, p0)


def test_threaded_oracle_equals_serial():
    g, q, fl, gas = _state()
    assert po.set_threads(1) == 1
    one = po.flux_div(g, gas, fl, q)
    assert po.set_threads(3) == 3
    three = po.flux_div(g, gas, fl, q)
    assert np.array_equal(one, three)


def test_cpu_baseline_reports_both_core_counts():
    cb = bench.cpu_baseline(sample_blocks=1, reps=2, max_threads=2, full_blocks=2)
    assert cb["kind"] == "port" and cb["unit"] == "Mcell-RHS/s"
    assert 1 <= cb["cores"] <= 2 and cb["value"] > 0.0
    assert cb["one_core"]["cores"] == 1 and cb["one_core"]["value"] > 0.0
    assert cb["full_grid"]["value"] > 0.0
