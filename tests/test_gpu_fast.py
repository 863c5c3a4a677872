"""GPU parity of the LDS-staged face-once flux_div kernel against the CPU oracle.

Blocks of 16x16xNz (the fast kernel's cross-section) with a rough,
seeded state; every combination the fast kernel serves.  Bar: bit-exact, or
1e-12 relative max-norm per variable where the power-law pow() enters.
"""
import numpy as np
import pytest

import golden_io as gi
from gpu_helpers import RawFlux
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

GAMMA, R = 1.4, 100.0 / 1.4
K2 = dict(kind=1, order=2)
VC = dict(kind=7, law=0, mu=1.0 / 1600.0, prandtl=0.72)
VP = dict(kind=7, law=1, mu=1.0 / 1600.0, T_ref=1.0, power=0.76, prandtl=0.72)
VP2 = dict(kind=7, law=1, mu=2.0e-3, T_ref=0.9, power=0.7, prandtl=0.7)
PDC = dict(kind=3, sensor=0, sensor_param=0.7, eps_p=0.3, eps_T=0.2)
PDD = dict(kind=3, sensor=1, sensor_param=1e-3, eps_p=0.3, eps_T=0.2)

COMBOS = {
    "keep2": [K2],
    "keep2_visc_const": [K2, VC],
    "keep2_visc_pow": [K2, VP],
    "keep2_visc_pow_tref": [K2, VP2],
    "keep2_pdiss_const": [K2, PDC],
    "keep2_pdiss_ducros_visc_pow": [K2, PDD, VP],
    "keep2_pdiss_const_visc_const": [K2, PDC, VC],
    "visc_pow": [VP],
    "pdiss_ducros": [PDD],
    "pdiss_ducros_visc_const": [PDD, VC],
}
POW = {k for k, v in COMBOS.items() if any(t.get("law") == 1 for t in v)}


def _state(nb, nc, ng, seed):
    rng = np.random.default_rng(seed)
    L = 2.0 * np.pi
    shape = (nb[0] * nb[1] * nb[2], nc[2] + 2 * ng, nc[1] + 2 * ng, nc[0] + 2 * ng, 5)
    q = np.zeros(shape)
    inner = q[:, ng:-ng, ng:-ng, ng:-ng]
    inner[..., 0] = R * (1.0 + 0.01 * rng.standard_normal(inner.shape[:-1]))
    inner[..., 1] = 1.0 + 0.02 * rng.standard_normal(inner.shape[:-1])
    for v in (2, 3, 4):
        inner[..., v] = 0.3 * rng.standard_normal(inner.shape[:-1])
    return q, L


def _run(nb, nc, terms, seed, periodic=(True, True, True), walls=(), increment=False, ng=2):
    import spade_amd as sa
    sa.load()
    host, L = _state(nb, nc, ng, seed)
    g = sa.CartesianGrid(nc, nb, ((0.0, L),) * 3, num_exch=ng, periodic=periodic)
    q = sa.GridArray(g).upload(host)
    sa.create_exchange(q).exchange(q)
    for face in walls:
        sa.boundary_fill_wall(q, face, 0.95)
    flux = RawFlux(terms, GAMMA, R)
    assert sa.flux_div_kernel(q, flux) == "fast"
    rhs = sa.GridArray(g, fill=0.0)
    init = None
    if increment:
        init = np.random.default_rng(seed + 1).standard_normal(g.host_shape)
        rhs.upload(init)
    sa.flux_div(q, rhs, flux, increment=increment)
    got = gi.interior(rhs.download(), ng)
    # oracle on exactly the ghosted state the GPU used
    og = po.make_grid(nb, nc, ng, (0, 0, 0), (L, L, L), periodic=periodic)
    oq = q.download()
    ofl = po.make_flux(*[po.make_term(**t) for t in terms])
    ref = po.flux_div(og, po.Gas(GAMMA, R), ofl, oq, init.copy() if increment else None, increment=increment)
    return got, gi.interior(ref, ng)


@pytest.mark.parametrize("name", sorted(COMBOS))
def test_fast_flux_div(name):
    got, ref = _run((2, 2, 2), (16, 16, 8), COMBOS[name], seed=7)
    err = gi.rel_maxnorm(got, ref)
    print(name, err)
    if name in POW:
        assert np.all(err <= 1e-12)
    else:
        assert np.array_equal(got, ref)


def test_fast_increment_cubic_blocks():
    got, ref = _run((2, 1, 2), (16, 16, 16), COMBOS["keep2_pdiss_const_visc_const"], seed=11, increment=True)
    assert np.array_equal(got, ref)


def test_fast_thin_blocks():
    got, ref = _run((1, 2, 3), (16, 16, 2), COMBOS["keep2_visc_const"], seed=5)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("nz, ng", [(1, 1), (3, 2), (5, 2)])
def test_fast_odd_and_single_plane_blocks(nz, ng):
    """The k-march takes planes in pairs: an odd number of planes ends on the
    first plane of a pair, and a single plane (one ghost layer: blocks are at
    least as wide as the ghost layers) is the prologue plus one step."""
    got, ref = _run((2, 1, 3), (16, 16, nz), COMBOS["keep2_pdiss_ducros_visc_pow"], seed=20 + nz,
                    increment=(nz == 3), ng=ng)
    err = gi.rel_maxnorm(got, ref)
    print("nz", nz, err, flush=True)
    assert np.all(err <= 1e-12), err


def test_fast_one_ghost_layer():
    """cent_keep<2>, pressure_diss_lr and visc_lr reach one cell across a
    face, so one ghost layer suffices: the x-halo then holds one section per
    side."""
    got, ref = _run((2, 2, 2), (16, 16, 3), COMBOS["keep2_pdiss_const_visc_const"], seed=31, ng=1)
    assert np.array_equal(got, ref)


def test_fast_walls():
    got, ref = _run((2, 2, 1), (16, 16, 4), COMBOS["keep2_pdiss_const_visc_const"], seed=3,
                    periodic=(True, False, True), walls=(2, 3))
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("increment", [False, True], ids=["overwrite", "increment"])
def test_fast_pow_outside_series_range(increment):
    """Hot and cold cells (T/T_ref far from 1) force the blocks that see them to
    leave the planes from the first affected one to the fix-up kernel; every
    block, and the increment mode, must still match the oracle."""
    import spade_amd as sa
    sa.load()
    nb, nc, ng = (2, 2, 2), (16, 16, 16), 2
    host, L = _state(nb, nc, ng, 11)
    inner = host[:, ng:-ng, ng:-ng, ng:-ng]
    inner[..., 1] = 1.0 + 0.005 * np.random.default_rng(3).standard_normal(inner.shape[:-1])
    inner[1, 5, 3, 4, 1] = 1.4       # block 1, plane 5
    inner[2, 0, 0, 0, 1] = 0.6       # block 2, plane 0 (and the ghost plane of its z-neighbour)
    inner[3, 15, 7, 15, 1] = 1.3     # block 3, last plane, x-boundary column
    inner[5, :, :, :, 1] = 1.25      # block 5 entirely out of range
    g = sa.CartesianGrid(nc, nb, ((0.0, L),) * 3, num_exch=ng)
    q = sa.GridArray(g).upload(host)
    sa.create_exchange(q).exchange(q)
    terms = [K2, VP]
    flux = RawFlux(terms, GAMMA, R)
    assert sa.flux_div_kernel(q, flux) == "fast"
    rhs = sa.GridArray(g, fill=0.0)
    init = None
    if increment:
        init = np.random.default_rng(5).standard_normal(g.host_shape)
        rhs.upload(init)
    sa.flux_div(q, rhs, flux, increment=increment)
    got = gi.interior(rhs.download(), ng)
    og = po.make_grid(nb, nc, ng, (0, 0, 0), (L, L, L))
    ref = po.flux_div(og, po.Gas(GAMMA, R), po.make_flux(*[po.make_term(**t) for t in terms]), q.download(),
                      init.copy() if increment else None, increment=increment)
    err = gi.rel_maxnorm(got, gi.interior(ref, ng))
    print("pow outside series range", err, flush=True)
    assert np.all(err <= 1e-12), err
