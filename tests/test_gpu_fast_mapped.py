"""The LDS-staged face-once flux_div kernel on mapped (diagonal, tanh)
coordinates, against the per-cell kernel of the same composition and against
the CPU oracle.

Grids of 16x16x16 blocks (the fast kernel's cross-section) stretched in y
(periodic in x and z, isothermal walls in y, as the channel), and stretched
along all three axes (every gradient component and face normal carries a map
factor); rough seeded states.  The mapped kernel evaluates each face once and
keeps the per-cell kernel's expression order, so it gives the per-cell kernel's bits wherever
no pow() enters (the fast kernel's power law is a series within one ulp of
pow, so 1e-12 there); against the oracle 1e-12 (the device tanh is OCML's,
the host's glibc's).  A variable that is zero everywhere (the viscous mass
flux) must be zero in both.
"""
import os

import numpy as np
import pytest

from gpu_helpers import RawFlux
from oracle import pyoracle as po
from test_gpu_fast import COMBOS, POW, GAMMA, R

pytestmark = pytest.mark.gpu

NB, NC, NG = (2, 2, 2), (16, 16, 16), 2
LO, HI = (0.0, -1.0, 0.0), (2.0, 1.0, 1.5)
TANH = (-1.0, 1.0, 1.5)
TOL = 1e-12


def _sa():
    import spade_amd as sa
    sa.load()
    return sa


def _setup(sa, all_axes):
    periodic = (True, False, True) if not all_axes else (False, False, False)
    lo = LO if not all_axes else (-1.0, -1.0, -1.0)
    hi = HI if not all_axes else (1.0, 1.0, 1.0)
    omaps = [None, (po.MAP_TANH,) + TANH, None] if not all_axes else [(po.MAP_TANH,) + TANH] * 3
    dmaps = tuple(None if m is None else sa.tanh_map(*m[1:]) for m in omaps)
    g = sa.CartesianGrid(NC, NB, tuple(zip(lo, hi)), num_exch=NG, periodic=periodic, maps=dmaps)
    og = po.make_grid(NB, NC, NG, lo, hi, periodic=[int(p) for p in periodic], maps=omaps)
    rng = np.random.default_rng(17)
    q = np.zeros(po.shape_of(og))
    inner = q[:, NG:-NG, NG:-NG, NG:-NG]
    n = inner.shape[:-1]
    inner[..., 0] = R * (1.0 + 0.01 * rng.standard_normal(n))
    inner[..., 1] = 1.0 + 0.02 * rng.standard_normal(n)
    for v in (2, 3, 4):
        inner[..., v] = 0.3 * rng.standard_normal(n)
    po.exchange(og, q)
    faces = (2, 3) if not all_axes else (0, 1, 2, 3, 4, 5)
    for face in faces:
        po.boundary_fill_wall(og, q, face, 0.95)
    return g, og, q


def _rel(a, b):
    """Relative max-norm per variable; a variable zero everywhere in b must be zero in a."""
    scale = np.max(np.abs(b), axis=0)
    diff = np.max(np.abs(a - b), axis=0)
    return np.where(scale > 0, diff / np.where(scale > 0, scale, 1.0), np.where(diff > 0, np.inf, 0.0))


def _flux_div(sa, q, terms, increment, disable_fast, init):
    rhs = sa.GridArray(q.grid).upload(init)
    if disable_fast:
        os.environ["SPD_DISABLE_FAST"] = "1"
    try:
        kind = sa.flux_div_kernel(q, RawFlux(terms, GAMMA, R))
        sa.flux_div(q, rhs, RawFlux(terms, GAMMA, R), increment=increment)
    finally:
        os.environ.pop("SPD_DISABLE_FAST", None)
    return kind, rhs.download()


@pytest.mark.parametrize("all_axes", [False, True], ids=["y_stretched", "xyz_stretched"])
@pytest.mark.parametrize("name", sorted(COMBOS))
def test_mapped_fast_equals_per_cell(name, all_axes):
    sa = _sa()
    g, og, host = _setup(sa, all_axes)
    q = sa.GridArray(g).upload(host)
    init = np.random.default_rng(3).standard_normal(g.host_shape)
    for increment in (False, True):
        kf, fast = _flux_div(sa, q, COMBOS[name], increment, False, init)
        ks, spec = _flux_div(sa, q, COMBOS[name], increment, True, init)
        assert (kf, ks) == ("fast", "specialized"), (kf, ks)
        sl = (slice(None), slice(NG, -NG), slice(NG, -NG), slice(NG, -NG))
        a, b = fast[sl].reshape(-1, 5), spec[sl].reshape(-1, 5)
        err = _rel(a, b)
        print(name, all_axes, increment, err, flush=True)
        if name in POW:
            assert np.all(err <= TOL), err
        else:
            assert np.array_equal(a, b), err


@pytest.mark.parametrize("name", ["keep2_pdiss_ducros_visc_pow", "keep2_pdiss_const_visc_const"])
def test_mapped_fast_matches_oracle(name):
    sa = _sa()
    g, og, host = _setup(sa, False)
    q = sa.GridArray(g).upload(host)
    rhs = sa.GridArray(g, fill=0.0)
    assert sa.flux_div_kernel(q, RawFlux(COMBOS[name], GAMMA, R)) == "fast"
    sa.flux_div(q, rhs, RawFlux(COMBOS[name], GAMMA, R), increment=False)
    ref = po.flux_div(og, po.Gas(GAMMA, R), po.make_flux(*[po.make_term(**t) for t in COMBOS[name]]), host)
    sl = (slice(None), slice(NG, -NG), slice(NG, -NG), slice(NG, -NG))
    a, b = rhs.download()[sl].reshape(-1, 5), ref[sl].reshape(-1, 5)
    err = _rel(a, b)
    print(name, "vs oracle", err)
    assert np.all(err <= TOL), err
