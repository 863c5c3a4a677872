"""The CPU-runnable Taylor-Green configuration at its own size, advanced over
several steps on the GPU and compared with the oracle after every step
count: 64^3 cells as 4x4x4 blocks of 16^3, two ghost layers, periodic
[0, 2 pi]^3, gamma 1.4, Mach 0.1, Re 1600, Pr 0.72, power-law viscosity
(n = 0.76), cent_keep<2> + visc_lr, dt from CFL 0.5 at t = 0 held fixed
(development/cuda-tgv/main.cc:155-248).

This is the path the 16^3 blocks take: the LDS-staged flux kernel (asserted)
and the stage kernels that write the exchanged ghosts themselves, over 10
steps of the transform-aware ssprk3_t (time-integration/advance.h:233-280)
and of the low-storage ssprk3_opt (advance.h:354-402).  Bar: 1e-12 relative
max-norm per variable (the device pow of the power law is within one ulp of
glibc's; everything else is the reference's rounding sequence).
"""
import numpy as np
import pytest

import bench
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

NB, NC, NG = (4, 4, 4), (16, 16, 16), 2
L = 2.0 * np.pi
GAMMA, MACH = 1.4, 0.1
R = (1.0 / MACH) ** 2 / GAMMA        # p0 = rho0 c0^2 / gamma with rho0 = u0 = T0 = 1
MU, T_REF, POWER, PR = 1.0 / 1600.0, 1.0, 0.76, 0.72
STEPS = 10
TOL = 1e-12


def _interior(a):
    return a[:, NG:-NG, NG:-NG, NG:-NG, :]


@pytest.mark.parametrize("scheme", ["ssprk3", "ssprk3_opt"])
def test_tgv64_ten_steps_match_oracle(scheme):
    import spade_amd as sa
    sa.load()
    g = sa.CartesianGrid(NC, NB, ((0.0, L),) * 3, num_exch=NG)
    host = np.zeros(g.host_shape)
    host[:, NG:-NG, NG:-NG, NG:-NG] = bench.tgv_interior(NB, NC, L, R)
    q = sa.GridArray(g).upload(host)
    sa.create_exchange(q).exchange(q)
    gas = sa.IdealGas(GAMMA, R)
    flux = sa.compose(sa.cent_keep(2, gas), sa.visc_lr(sa.power_law(MU, T_REF, POWER, PR), gas))
    assert sa.flux_div_kernel(q, flux) == "fast"
    dt = 0.5 * (L / 64) / sa.max_wavespeed(q, gas)

    og = po.make_grid(NB, NC, NG, (0, 0, 0), (L, L, L))
    ofl = po.make_flux(po.make_term(po.TERM_KEEP, order=2),
                       po.make_term(po.TERM_VISC, law=po.LAW_POWER, mu=MU, T_ref=T_REF, power=POWER, prandtl=PR))
    oq = q.download()
    step = po.ssprk3_step if scheme == "ssprk3" else po.ssprk3opt_step
    it = sa.Integrator(q, flux, dt, scheme=getattr(sa, scheme))
    for n in range(1, STEPS + 1):
        it.advance(1)
        step(og, po.Gas(GAMMA, R), ofl, None, dt, oq, 1)
        got = q.download()
        err = np.max(np.abs(_interior(got) - _interior(oq)).reshape(-1, 5), axis=0) / \
            np.max(np.abs(_interior(oq)).reshape(-1, 5), axis=0)
        print(f"{scheme} step {n}: rel err per variable {err}", flush=True)
        assert np.all(err <= TOL), (n, err)
        # the fused ghosts equal what the stand-alone exchange copies
        if n == STEPS:
            sa.create_exchange(q).exchange(q)
            assert np.array_equal(q.download(), got)
