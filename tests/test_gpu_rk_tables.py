"""Explicit Runge-Kutta with any Butcher table, caller-supplied residual and
boundary callbacks, and the stand-alone stage update, on the GPU through the
C ABI, against golden vectors of the compiled reference (G2: the 16^3 TGV as
2x2x2 blocks of 8^3, cent_keep<2> + visc_lr power law, dt from CFL 0.5).

* transform-aware advance (time-integration/advance.h:233-280) with rk2_t,
  rk4_t, ssprk34_t and rk38r_t, primitive storage, two steps;
* generic advance (advance.h:107-230) with the identity transform on
  conserved storage: the residual callback converts every stored cell to
  primitives and takes flux_div, the boundary callback exchanges; rk4_t and
  rk38r_t with one solution copy, rk2hs_t and ssprk3hs_t with two;
* spd_rk_stage: one SSP-RK3 step assembled from spd_flux_div, spd_rk_stage
  and spd_exchange calls equals the reference's step.

Bar: 1e-12 relative max-norm per variable (the power law's pow is the device's).
"""
import numpy as np
import pytest

import golden_io as gi
from cases import FLUX_CASES, G2, GAMMA, R, P
from gpu_helpers import RawFlux, grid

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _sa():
    import spade_amd as sa
    sa.load()
    return sa


def _prim_state(sa):
    g = grid(G2)
    q = sa.GridArray(g).upload(gi.embed(gi.load("g2_prim0"), 2))
    sa.create_exchange(q).exchange(q)
    return g, q


def _flux(sa):
    return RawFlux(FLUX_CASES["keep2_visc_pow"], GAMMA, R)


@pytest.mark.parametrize("name, table", [("rk2", "rk2"), ("rk4", "rk4"), ("ssprk34", "ssprk34"), ("rk38r", "rk38r")])
def test_transform_aware_tables(name, table):
    sa = _sa()
    g, q = _prim_state(sa)
    it = sa.RKIntegrator(q, getattr(sa, table), P["g2_dt"], trans=sa.prim_cons, gas=sa.IdealGas(GAMMA, R),
                         flux=_flux(sa))
    it.advance(2)
    err = gi.rel_maxnorm(gi.interior(q.download(), 2), gi.load(f"g2_{name}_step2"))
    print(name, err)
    assert np.all(err <= TOL), err
    assert it.time() == _axis_time(sa.rk_table(getattr(sa, table)), P["g2_dt"], 2)


def _axis_time(table, dt, steps):
    """time_axis_t after `steps` advances: each stage moves the time to t + c dt
    and back, the step then adds dt (advance.h:177-182, 268-279), in doubles."""
    t = 0.0
    for _ in range(steps):
        for i in range(table.stages):
            c = table.c[i]
            tc = 0.0 if c.den == 0 else float(c.num) / float(c.den)
            t += tc * dt
            t -= tc * dt
        t += dt
    return t


def test_ssprk3_table_equals_fixed_integrator():
    """The general path with the ssprk3_t table gives the fixed integrator's bits
    (same stage arithmetic; the general path exchanges after the stage instead
    of writing the ghosts from the stage kernel)."""
    sa = _sa()
    g, q1 = _prim_state(sa)
    _, q2 = _prim_state(sa)
    sa.Integrator(q1, _flux(sa), P["g2_dt"], scheme=sa.ssprk3).advance(3)
    sa.RKIntegrator(q2, sa.ssprk3_table, P["g2_dt"], gas=sa.IdealGas(GAMMA, R), flux=_flux(sa)).advance(3)
    assert np.array_equal(q1.download(), q2.download())


@pytest.mark.parametrize("table", ["rk4", "rk38r", "rk2hs", "ssprk3hs"])
def test_identity_transform_conserved_storage(table):
    sa = _sa()
    g = grid(G2)
    gas = sa.IdealGas(GAMMA, R)
    q = sa.GridArray(g).upload(gi.load("g2c_cons0"))
    scratch = sa.GridArray(g)
    flux = _flux(sa)
    times = []

    def rhs(resid, sol, t):
        scratch.copy_from(sol)
        scratch.convert("prim", gas, include_ghosts=True)
        sa.flux_div(scratch, resid, flux, increment=False)
        times.append(t)

    def bc(sol, t):
        sa.create_exchange(sol).exchange(sol)

    it = sa.RKIntegrator(q, getattr(sa, table), P["g2_dt"], trans=sa.identity_transform, rhs=rhs, bc=bc)
    it.advance(2)
    err = gi.rel_maxnorm(q.download(), gi.load(f"g2c_{table}_step2"))
    print(table, err)
    assert np.all(err <= TOL), err
    assert len(times) == 2 * sa.rk_table(getattr(sa, table)).stages


def test_callback_error_is_reported():
    sa = _sa()
    g, q = _prim_state(sa)

    def rhs(resid, sol, t):
        raise ValueError("planted")

    it = sa.RKIntegrator(q, sa.rk4, P["g2_dt"], trans=sa.identity_transform, rhs=rhs)
    with pytest.raises(ValueError, match="planted"):
        it.advance(1)


def test_rk_stage_assembles_the_reference_step():
    """One transform-aware SSP-RK3 step from stand-alone calls: residual k0,
    stage (1)dt, exchange, k1, stage (-3/4, 1/4)dt, exchange, k2, stage
    (-1/12, -1/12, 2/3)dt, exchange (advance.h:233-280 with the ssprk3_t rows)."""
    sa = _sa()
    g, q = _prim_state(sa)
    gas = sa.IdealGas(GAMMA, R)
    flux = _flux(sa)
    dt = P["g2_dt"]
    k = [sa.GridArray(g, fill=0.0) for _ in range(3)]
    ex = sa.create_exchange(q)
    sa.flux_div(q, k[0], flux, increment=False)
    sa.rk_stage(q, k[:1], [(1.0 / 1.0) * dt], gas)
    ex.exchange(q)
    sa.flux_div(q, k[1], flux, increment=False)
    sa.rk_stage(q, k[:2], [(-3.0 / 4.0) * dt, (1.0 / 4.0) * dt], gas)
    ex.exchange(q)
    sa.flux_div(q, k[2], flux, increment=False)
    sa.rk_stage(q, k, [(-1.0 / 12.0) * dt, (-1.0 / 12.0) * dt, (2.0 / 3.0) * dt], gas, push_ghosts=True)
    pushed = q.download()
    ex.exchange(q)
    got = q.download()
    assert np.array_equal(pushed, got)          # the pushed ghosts are what the exchange copies
    err = gi.rel_maxnorm(gi.interior(got, 2), gi.load("g2_ssprk3_step1"))
    assert np.all(err <= TOL), err


def test_array_operations():
    sa = _sa()
    g, q = _prim_state(sa)
    host = q.download()
    other = sa.GridArray(g).upload(host * 0.5 + 1.0)
    q += other
    want = host.copy()
    want[:, 2:-2, 2:-2, 2:-2] += (host * 0.5 + 1.0)[:, 2:-2, 2:-2, 2:-2]   # interior only
    assert np.array_equal(q.download(), want)
    q -= other
    want[:, 2:-2, 2:-2, 2:-2] -= (host * 0.5 + 1.0)[:, 2:-2, 2:-2, 2:-2]
    assert np.array_equal(q.download(), want)
    q *= 0.375
    assert np.array_equal(q.download(), want * 0.375)                       # every stored cell


def test_bad_tables_refused():
    sa = _sa()
    g, q = _prim_state(sa)
    t = sa.rk_table(sa.rk4)
    t.a[1][2].num, t.a[1][2].den = 1, 2          # not explicit
    with pytest.raises(sa.SpadeError):
        sa.RKIntegrator(q, t, P["g2_dt"], gas=sa.IdealGas(GAMMA, R), flux=_flux(sa))
    t = sa.rk_table(sa.rk4)
    t.stages = 7
    with pytest.raises(sa.SpadeError):
        sa.RKIntegrator(q, t, P["g2_dt"], gas=sa.IdealGas(GAMMA, R), flux=_flux(sa))
    with pytest.raises(sa.SpadeError):          # the default residual needs primitive storage
        sa.RKIntegrator(q, sa.rk4, P["g2_dt"], trans=sa.identity_transform, flux=_flux(sa))
