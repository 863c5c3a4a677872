"""The wall-bounded curvilinear channel at its full size (192 x 128 x 96 cells,
12 x 8 x 6 blocks of 16^3, tanh wall-normal stretching of factor 1.5,
isothermal no-slip walls, Mach 0.2, cent_keep<2> + pressure_diss_lr + visc_lr
with the power law; spade_amd.studies.channel).

* one SSP-RK3 step on the GPU equals the CPU oracle's step within 1e-12
  relative max-norm per variable (device tanh / pow against glibc's);
* 200 steps: the state stays finite and bounded, and without the pressure
  diffusion term (whose eps_T part moves mass through a wall whose ghost
  temperature is mirrored) the total mass sum rho x' y' z' is conserved to
  round-off.
"""
import math

import numpy as np
import pytest

import golden_io as gi
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _setup(with_pdiss=True):
    import spade_amd as sa
    from spade_amd import studies
    sa.load()
    c = studies.channel()
    periodic = (True, False, True)
    g = sa.CartesianGrid(c["nc"], c["nb"], tuple(zip(c["lo"], c["hi"])), num_exch=c["ng"], periodic=periodic,
                         maps=(None, sa.tanh_map(*c["tanh_y"]), None))
    gas = sa.IdealGas(c["gamma"], c["R"])
    law = sa.power_law(c["mu"], 1.0, c["power"], c["prandtl"])
    terms = [sa.cent_keep(2, gas)]
    oterms = [po.make_term(po.TERM_KEEP, order=2)]
    if with_pdiss:
        a, ep, et = c["pdiss"]
        terms.append(sa.pressure_diss_lr(gas, sa.const_sensor(a), ep, et))
        oterms.append(po.make_term(po.TERM_PDISS, sensor=po.SENSOR_CONST, sensor_param=a, eps_p=ep, eps_T=et))
    terms.append(sa.visc_lr(law, gas))
    oterms.append(po.make_term(po.TERM_VISC, law=po.LAW_POWER, mu=c["mu"], T_ref=1.0, power=c["power"],
                               prandtl=c["prandtl"]))
    og = po.make_grid(c["nb"], c["nc"], c["ng"], c["lo"], c["hi"], periodic=(1, 0, 1),
                      maps=(None, (po.MAP_TANH,) + tuple(c["tanh_y"]), None))
    return sa, c, g, gas, sa.compose(*terms), og, po.make_flux(*oterms)


def _mass(prims, c, R):
    """sum of rho x' y' z' over the interior (x' and z' are constant)."""
    nb, ng = c["nb"], c["ng"]
    lo, hi = c["lo"], c["hi"]
    beta = c["tanh_y"][2]
    dxi = (hi[1] - lo[1]) / (nb[1] * 16)
    inner = prims[:, ng:-ng, ng:-ng, ng:-ng]
    rho = inner[..., 0] / (R * inner[..., 1])
    total = 0.0
    for lb in range(rho.shape[0]):
        by = (lb // nb[0]) % nb[1]
        xi = lo[1] + (by * 16 + np.arange(16) + 0.5) * dxi
        t = np.tanh(beta * xi)
        yp = beta * (1 - t * t) / math.tanh(beta)
        total += float(np.sum(rho[lb] * yp[None, :, None]))
    return total


def test_channel_step_matches_oracle():
    sa, c, g, gas, flux, og, oflux = _setup()
    print("channel: 192x128x96, 1 step, dt", c["dt"], flush=True)
    start = c["host"].copy()               # boundary conditions applied before the first step
    po.exchange(og, start)
    po.boundary_fill_wall(og, start, 2, c["T_wall"])
    po.boundary_fill_wall(og, start, 3, c["T_wall"])
    q = sa.GridArray(g).upload(start)
    it = sa.Integrator(q, flux, c["dt"], scheme=sa.ssprk3, walls=(2, 3), t_wall=c["T_wall"])
    it.advance(1)
    got = gi.interior(q.download(), c["ng"])
    print("channel: GPU step done, oracle step on the host", flush=True)
    ref = start
    bc = po.BC(wall_mask=(1 << 2) | (1 << 3), t_wall=c["T_wall"])
    po.ssprk3_step(og, po.Gas(c["gamma"], c["R"]), oflux, bc, c["dt"], ref)
    print("channel: oracle step done", flush=True)
    err = gi.rel_maxnorm(got, gi.interior(ref, c["ng"]))
    print("channel 1 step rel err per var", err, flush=True)
    assert np.all(err <= TOL), err


@pytest.mark.parametrize("with_pdiss", [False, True], ids=["keep2_visc", "keep2_pdiss_visc"])
def test_channel_200_steps(with_pdiss):
    sa, c, g, gas, flux, og, oflux = _setup(with_pdiss)
    print("channel: 200 steps, pdiss", with_pdiss, flush=True)
    q = sa.GridArray(g).upload(c["host"])
    sa.create_exchange(q).exchange(q)
    sa.boundary_fill_wall(q, 2, c["T_wall"])
    sa.boundary_fill_wall(q, 3, c["T_wall"])
    it = sa.Integrator(q, flux, c["dt"], scheme=sa.ssprk3, walls=(2, 3), t_wall=c["T_wall"])
    # mass of the initial state (ghosts are not part of the sum)
    m0 = _mass(c["host"], c, c["R"])
    for chunk in range(4):
        it.advance(50)
        g.ctx.sync()
        print("channel: steps", 50 * (chunk + 1), flush=True)
    out = q.download()
    inner = gi.interior(out, c["ng"])
    assert np.all(np.isfinite(inner))
    u = inner[..., 2]
    print("channel after 200 steps: max|u|", float(np.max(np.abs(u))), "T range",
          float(inner[..., 1].min()), float(inner[..., 1].max()), flush=True)
    assert float(np.max(np.abs(u))) < 2.0
    assert 0.9 < float(inner[..., 1].min()) and float(inner[..., 1].max()) < 1.1
    drift = abs(_mass(out, c, c["R"]) - m0) / m0
    print("relative mass drift", drift, flush=True)
    if not with_pdiss:
        assert drift <= 1e-12, drift
    else:
        assert drift <= 1e-6, drift
