"""The headline workload at its full size on the GPU: FP64 flux_div
(cent_keep<2> + visc_lr with the power law) on the 256^3 Taylor-Green vortex,
16^3 blocks of 16^3 cells, two ghost layers -- the state bench.py times.

* eight blocks (the corners of the 16^3 block grid and two inside it) equal
  the oracle evaluated on the same blocks of the same array, within 1e-12
  relative max-norm per variable (the device pow of the power law is within
  one ulp of glibc's, everything else bit for bit);
* discrete conservation over the whole grid: every face flux enters the two
  cells it separates with opposite signs, so each residual variable sums to
  zero up to the rounding of the per-cell sums;
* one SSP-RK3 step of the same state conserves total mass (sum of p / (R T)
  over the interior) to round-off;
* two SSP-RK3 steps in one advance() call through the fused stages (their
  halos read from the neighbour blocks, workgroups in runs of blocks per XCD)
  equal the separate flux_div and stage launches bit for bit, ghost cells
  included, over all 4096 blocks.
"""
import os

import numpy as np
import pytest

import bench
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

NB, NC, NG = (16, 16, 16), (16, 16, 16), 2
L = 2.0 * np.pi
GAMMA, R = 1.4, 100.0 / 1.4
MU, T_REF, POWER, PR = 1.0 / 1600.0, 1.0, 0.76, 0.72
BLOCKS = [0, 15, 240, 255, 3840, 3855, 4080, 4095, 1911, 2184]


def _setup():
    import spade_amd as sa
    sa.load()
    g = sa.CartesianGrid(NC, NB, ((0.0, L),) * 3, num_exch=NG)
    host = np.zeros(g.host_shape)
    host[:, NG:-NG, NG:-NG, NG:-NG] = bench.tgv_interior(NB, NC, L, R)
    q = sa.GridArray(g).upload(host)
    sa.create_exchange(q).exchange(q)
    gas = sa.IdealGas(GAMMA, R)
    flux = sa.compose(sa.cent_keep(2, gas), sa.visc_lr(sa.power_law(MU, T_REF, POWER, PR), gas))
    return sa, g, q, gas, flux


def _interior(a):
    return a[:, NG:-NG, NG:-NG, NG:-NG, :]


def test_flux_div_256_blocks_match_oracle_and_conserve():
    sa, g, q, gas, flux = _setup()
    rhs = sa.GridArray(g, fill=0.0)
    sa.flux_div(q, rhs, flux, increment=False)
    got = _interior(rhs.download())
    prim = q.download()
    print("256^3: GPU flux_div done, oracle on", len(BLOCKS), "blocks", flush=True)
    og = po.make_grid(NB, NC, NG, (0, 0, 0), (L, L, L))
    ofl = po.make_flux(po.make_term(po.TERM_KEEP, order=2),
                       po.make_term(po.TERM_VISC, law=po.LAW_POWER, mu=MU, T_ref=T_REF, power=POWER, prandtl=PR))
    ref = _interior(po.flux_div_blocks(og, po.Gas(GAMMA, R), ofl, prim, BLOCKS))
    for lb in BLOCKS:
        scale = np.max(np.abs(ref[lb].reshape(-1, 5)), axis=0)
        err = np.max(np.abs((got[lb] - ref[lb]).reshape(-1, 5)), axis=0) / scale
        assert np.all(err <= 1e-12), (lb, err)
    total = got.reshape(-1, 5).sum(axis=0)
    mag = np.abs(got).reshape(-1, 5).sum(axis=0)
    print("256^3 residual sums relative to sum |rhs|:", total / mag, flush=True)
    assert np.all(np.abs(total) <= 1e-12 * mag), total / mag


def test_ssprk3_step_256_conserves_mass():
    sa, g, q, gas, flux = _setup()

    def mass():
        p = _interior(q.download())
        return float(np.sum(p[..., 0] / (R * p[..., 1])))

    m0 = mass()
    dt = 0.5 * (L / 256) / sa.max_wavespeed(q, gas)
    sa.Integrator(q, flux, dt, scheme=sa.ssprk3).advance(1)
    m1 = mass()
    print("256^3 one step: relative mass change", (m1 - m0) / m0, flush=True)
    assert abs(m1 - m0) <= 1e-12 * m0


def test_fused_ssprk3_256_equals_separate_launches():
    def run(fused):
        if fused:
            os.environ.pop("SPD_DISABLE_FUSED", None)
        else:
            os.environ["SPD_DISABLE_FUSED"] = "1"
        try:
            sa, g, q, gas, flux = _setup()
            dt = 0.5 * (L / 256) / sa.max_wavespeed(q, gas)
            it = sa.Integrator(q, flux, dt, scheme=sa.ssprk3)
            assert it.fused == fused
            it.advance(2)
            out = q.download()
            it.close()
            return out
        finally:
            os.environ.pop("SPD_DISABLE_FUSED", None)
    a = run(True)
    print("256^3 two fused steps done", flush=True)
    b = run(False)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)
