"""The integrator's stage kernels write the exchanged ghost cells themselves
(the exchange of each stage result is fused into the stage update).  After
any number of steps the ghosts must equal, bit for bit, what the stand-alone
exchange (grid/exchange_pattern.h:382-424) followed by the wall fill copies
from the same interior -- on every stored cell, for:

* a periodic multi-block grid (both SSP-RK3 forms);
* a grid with isothermal walls on the non-periodic y axis (edges and corners
  that cross a wall are not exchanged; the wall fill runs after);
* one block along an axis (the block is its own neighbour) and blocks no wider
  than the ghost layers (a cell feeds the low and the high ghosts at once);
* one ghost layer (walls, and 16x16 blocks through the LDS-staged flux kernel);
* 16^3 blocks (the LDS-staged flux kernel), several steps.

The stand-alone exchange itself takes its workgroups in runs of blocks per
XCD by default; SPD_EXCH_XCD=0 gives the launch order, and both give the same
array on grids whose workgroup counts are and are not multiples of eight, with
periodic and non-periodic axes.
"""
import numpy as np
import pytest

import golden_io as gi
from cases import FLUX_CASES, GAMMA, R, P
from gpu_helpers import RawFlux

pytestmark = pytest.mark.gpu


def _sa():
    import spade_amd as sa
    sa.load()
    return sa


def _state(sa, g, seed):
    """A smooth state with a little noise (T near 1) and current ghosts."""
    rng = np.random.default_rng(seed)
    host = np.zeros(g.host_shape)
    ng = g.num_exch
    inner = host[:, ng:-ng, ng:-ng, ng:-ng]
    inner[..., 0] = P["p0"] * (1.0 + 0.01 * rng.standard_normal(inner.shape[:-1]))
    inner[..., 1] = 1.0 + 0.01 * rng.standard_normal(inner.shape[:-1])
    inner[..., 2:] = 0.1 * rng.standard_normal(inner.shape[:-1] + (3,))
    q = sa.GridArray(g).upload(host)
    sa.create_exchange(q).exchange(q)
    return q


def _check(sa, g, scheme, steps, walls=(), seed=1, dt=1e-4):
    q = _state(sa, g, seed)
    for f in walls:
        sa.boundary_fill_wall(q, f, P["t_wall"])
    it = sa.Integrator(q, RawFlux(FLUX_CASES["keep2_visc_pow"], GAMMA, R), dt, scheme=scheme,
                       walls=walls, t_wall=P["t_wall"])
    it.advance(steps)
    fused = q.download()
    sa.create_exchange(q).exchange(q)
    for f in walls:
        sa.boundary_fill_wall(q, f, P["t_wall"])
    pulled = q.download()
    assert np.all(np.isfinite(fused))
    assert np.array_equal(fused, pulled), np.argwhere(fused != pulled)[:5]


@pytest.mark.parametrize("scheme", ["ssprk3", "ssprk3_opt"])
def test_periodic(scheme):
    sa = _sa()
    g = sa.CartesianGrid((4, 8, 4), (2, 3, 2), ((0.0, 1.0), (0.0, 2.0), (0.0, 1.5)), num_exch=2)
    _check(sa, g, getattr(sa, scheme), 3)


def test_walls():
    sa = _sa()
    g = sa.CartesianGrid((4, 8, 4), (2, 3, 2), ((0.0, 1.0), (0.0, 2.0), (0.0, 1.5)), num_exch=2,
                         periodic=(True, False, True))
    _check(sa, g, sa.ssprk3, 3, walls=(2, 3))


def test_self_neighbour_and_narrow_blocks():
    sa = _sa()
    g = sa.CartesianGrid((2, 3, 4), (1, 2, 3), ((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)), num_exch=2)
    _check(sa, g, sa.ssprk3, 2)
    _check(sa, g, sa.ssprk3_opt, 2, seed=2)


def test_one_ghost_layer():
    sa = _sa()
    g = sa.CartesianGrid((4, 8, 3), (2, 3, 2), ((0.0, 1.0), (0.0, 2.0), (0.0, 1.5)), num_exch=1,
                         periodic=(True, False, True))
    _check(sa, g, sa.ssprk3, 3, walls=(2, 3))
    g = sa.CartesianGrid((16, 16, 3), (2, 1, 2), ((0.0, 1.0),) * 3, num_exch=1)
    _check(sa, g, sa.ssprk3_opt, 2, seed=4)


def test_fast_kernel_blocks():
    sa = _sa()
    g = sa.CartesianGrid((16, 16, 16), (2, 2, 2), ((0.0, 1.0),) * 3, num_exch=2)
    _check(sa, g, sa.ssprk3, 4)


@pytest.mark.parametrize("nb,nc,periodic", [((3, 5, 2), (16, 16, 16), (True, False, True)),
                                            ((4, 4, 4), (8, 12, 16), (True, True, True)),
                                            ((5, 1, 3), (6, 6, 6), (False, True, False))],
                         ids=["walls", "periodic", "thin"])
def test_exchange_xcd_runs_equal_launch_order(nb, nc, periodic):
    import os
    sa = _sa()
    g = sa.CartesianGrid(nc, nb, ((0.0, 1.0),) * 3, num_exch=2, periodic=periodic)
    host = np.random.default_rng(17).standard_normal(g.host_shape)
    print("exchange xcd runs", nb, nc, periodic, flush=True)
    out = {}
    for x in ("0", "1"):
        os.environ["SPD_EXCH_XCD"] = x
        try:
            q = sa.GridArray(g).upload(host)
            sa.create_exchange(q).exchange(q)
            out[x] = q.download()
        finally:
            os.environ.pop("SPD_EXCH_XCD", None)
    assert not np.array_equal(out["0"], host)
    assert np.array_equal(out["0"], out["1"])
