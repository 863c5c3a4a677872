"""The CPU oracle against the golden vectors of the compiled reference.

Every fixture in tests/golden was written by oracle/ref/ref_driver.cpp, built
against the SPADE headers themselves.  The oracle must reproduce each one bit
for bit (same expression order, same libm, no contraction).
"""
import numpy as np
import pytest

import golden_io as gi
from cases import FLUX_CASES, G1, G2, GAMMA, R, P
from oracle import pyoracle as po

T = gi.TWO_PI_REF


def _grid(spec, periodic=(1, 1, 1)):
    return po.make_grid(spec["nb"], spec["nc"], spec["ng"], (0, 0, 0), (T, T, T), periodic=periodic)


def _flux(terms):
    return po.make_flux(*[po.make_term(**t) for t in terms])


@pytest.fixture(scope="module")
def g1_prim():
    g = _grid(G1)
    a = gi.embed(gi.load("g1_prim"), 2)
    po.exchange(g, a)
    return a


def test_periodic_exchange_bitwise():
    g = _grid(G1)
    a = gi.embed(gi.load("g1_prim"), 2)
    po.exchange(g, a)
    assert np.array_equal(a, gi.load("g1_prim_ghost"))


@pytest.mark.parametrize("name", sorted(FLUX_CASES))
def test_flux_div_bitwise(name, g1_prim):
    g = _grid(G1)
    rhs = po.flux_div(g, po.Gas(GAMMA, R), _flux(FLUX_CASES[name]), g1_prim)
    assert np.array_equal(gi.interior(rhs, 2), gi.load("g1_rhs_" + name))


def test_flux_div_increment_bitwise(g1_prim):
    g = _grid(G1)
    init = gi.embed(gi.load("g1_rhs_incr_init"), 2)
    rhs = po.flux_div(g, po.Gas(GAMMA, R), _flux([dict(kind=1, order=2)]), g1_prim, init, increment=True)
    assert np.array_equal(gi.interior(rhs, 2), gi.load("g1_rhs_incr_keep2"))


def test_wall_exchange_and_boundary_fill_bitwise():
    g = _grid(G1, periodic=(1, 0, 1))
    a = gi.load("g3_prim_init_ghost").copy()
    po.exchange(g, a)
    assert np.array_equal(a, gi.load("g3_prim_exch_ghost"))
    po.boundary_fill_wall(g, a, 2, P["t_wall"])
    po.boundary_fill_wall(g, a, 3, P["t_wall"])
    assert np.array_equal(a, gi.load("g3_prim_wall_ghost"))
    rhs = po.flux_div(g, po.Gas(GAMMA, R), _flux([dict(kind=1, order=2), dict(kind=7, law=0, mu=P["mu_ref"],
                                                  prandtl=P["prandtl"])]), a)
    assert np.array_equal(gi.interior(rhs, 2), gi.load("g3_rhs_keep2_visc_const"))


ALL_FACES = 0b111111


def _g4():
    """The non-periodic grid after its exchange (the state the fills start from)."""
    g = _grid(G1, periodic=(0, 0, 0))
    a = gi.load("g3_prim_init_ghost").copy()       # same initial values as the G3 grid
    po.exchange(g, a)
    assert np.array_equal(a, gi.load("g4_prim_exch_ghost"))
    return g, a


@pytest.mark.parametrize("order, mask, name", [(2, ALL_FACES, "g4_prim_extrap2_ghost"),
                                                (0, ALL_FACES, "g4_prim_extrap0_ghost"),
                                                (3, 0b001001, "g4_prim_extrap3_xmin_ymax_ghost")])
def test_boundary_fill_extrapolation_bitwise(order, mask, name):
    g, a = _g4()
    po.boundary_fill_extrap(g, a, mask, order)
    assert np.array_equal(a, gi.load(name))


def slip_kernel(t_wall):
    """The reference driver's four-argument slip-wall kernel as an affine
    kernel: p = 0.25 p_ghost + p_image, T = 2 T_wall - T_image, normal
    velocity mirrored, tangential velocity copied."""
    return po.fill_kernel(dom=(1.0, -1.0, 1.0, 1.0, 1.0), fill=(0.25, 0.0, 0.0, 0.0, 0.0),
                          add=(0.0, 2.0 * t_wall, 0.0, 0.0, 0.0), dom_normal=-1.0, fill_normal=0.0, add_normal=0.0)


def test_boundary_fill_user_kernel_bitwise():
    g, a = _g4()
    po.boundary_fill_affine(g, a, ALL_FACES, slip_kernel(P["t_wall"]))
    assert np.array_equal(a, gi.load("g4_prim_slip_ghost"))


def test_boundary_fill_wall_kernel_is_affine():
    """The isothermal no-slip mirror written as an affine kernel gives the
    fixed wall fill's bits."""
    g = _grid(G1, periodic=(1, 0, 1))
    a = gi.load("g3_prim_exch_ghost").copy()
    k = po.fill_kernel(dom=(1.0, -1.0, -1.0, -1.0, -1.0), add=(0.0, 2.0 * P["t_wall"], 0.0, 0.0, 0.0))
    po.boundary_fill_affine(g, a, 0b001100, k)
    assert np.array_equal(a, gi.load("g3_prim_wall_ghost"))


def _g2_setup():
    g = _grid(G2)
    q = gi.embed(gi.load("g2_prim0"), 2)
    po.exchange(g, q)
    fl = _flux(FLUX_CASES["keep2_visc_pow"])
    return g, q, fl


def test_tgv_rhs_bitwise():
    g, q, fl = _g2_setup()
    rhs = po.flux_div(g, po.Gas(GAMMA, R), fl, q)
    assert np.array_equal(gi.interior(rhs, 2), gi.load("g2_rhs0"))


@pytest.mark.parametrize("steps,name", [(1, "g2_ssprk3_step1"), (10, "g2_ssprk3_step10")])
def test_ssprk3_bitwise(steps, name):
    g, q, fl = _g2_setup()
    po.ssprk3_step(g, po.Gas(GAMMA, R), fl, po.BC(0, 0.0), P["g2_dt"], q, steps)
    assert np.array_equal(gi.interior(q, 2), gi.load(name))


def test_ssprk3_opt_bitwise():
    g, q, fl = _g2_setup()
    po.ssprk3opt_step(g, po.Gas(GAMMA, R), fl, po.BC(0, 0.0), P["g2_dt"], q, 3)
    assert np.array_equal(gi.interior(q, 2), gi.load("g2_ssprk3opt_step3"))


def test_golden_checksums():
    """The fixtures are the files make_golden.sh wrote (sha256 manifest)."""
    import hashlib
    import os
    with open(os.path.join(gi.GOLDEN, "SHA256SUMS")) as fh:
        for line in fh:
            digest, name = line.split()
            with open(os.path.join(gi.GOLDEN, name), "rb") as f:
                assert hashlib.sha256(f.read()).hexdigest() == digest, name


@pytest.mark.parametrize("table", ["rk2", "rk4", "ssprk34", "rk38r"])
def test_transform_aware_tables_bitwise(table):
    """Two steps of the transform-aware advance (time-integration/advance.h:
    233-280) with the other Butcher tables of explicit.h, primitive storage."""
    g, q, fl = _g2_setup()
    po.rk_transform_step(g, po.Gas(GAMMA, R), fl, None, po.rk_table(table), P["g2_dt"], q, steps=2)
    assert np.array_equal(gi.interior(q, 2), gi.load(f"g2_{table}_step2"))


def test_transform_aware_ssprk3_table_equals_fixed_form():
    """The general-table path with the ssprk3_t table is the fixed ssprk3 step."""
    g, q, fl = _g2_setup()
    po.rk_transform_step(g, po.Gas(GAMMA, R), fl, None, po.rk_table("ssprk3"), P["g2_dt"], q, steps=1)
    assert np.array_equal(gi.interior(q, 2), gi.load("g2_ssprk3_step1"))


@pytest.mark.parametrize("table", ["rk4", "rk38r", "rk2hs", "ssprk3hs"])
def test_identity_transform_tables_bitwise(table):
    """Two steps of the generic advance (advance.h:107-230) with the identity
    transform on conserved storage (residual: primitives of every stored cell,
    then flux_div; boundary: exchange), one and two solution copies."""
    g = _grid(G2)
    q = gi.load("g2c_cons0").copy()
    fl = _flux(FLUX_CASES["keep2_visc_pow"])
    po.rk_identity_step(g, po.Gas(GAMMA, R), fl, po.rk_table(table), P["g2_dt"], q, steps=2)
    assert np.array_equal(q, gi.load(f"g2c_{table}_step2"))
