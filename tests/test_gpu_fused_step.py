"""The SSP-RK3 step with each stage fused into one flux_div launch equals the
separate flux_div + stage launches bit for bit, ghost cells included
(time-integration/advance.h:56-102 applied to rhs = flux_div).

The fused stage writes its result into a second buffer (the integrator's)
and the two buffers trade places after each step.  On periodic grids, and
on the mapped channel (periodic in x and z, isothermal walls in y), a stage
after the first of an advance() call reads its halo from the neighbour
blocks' interiors (forming the wall ghosts itself) and only the last stage
writes the ghost cells and fills the walls; elsewhere every stage pushes the
ghosts its result feeds.  The cases cover one and several
steps per call (the state alternating between the buffers), a periodic
Taylor-Green grid, blocks of 16 x 16 x 8 and block rows that are their own
neighbours, ghost cells the caller left unexchanged (read by the first
stage, as the separate launches do), a state whose face temperatures leave
the power-law series range in some blocks (those planes are completed by the
fix-up kernels of the fused launch, after their blocks' ghosts are pulled),
and the stretched channel with isothermal walls (one and several steps per
call, and blocks at both walls left to the fix-up).
SPD_DISABLE_FUSED=1 selects the separate launches.
"""
import os

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu

L = 2.0 * np.pi
GAMMA, MACH = 1.4, 0.1
R = (1.0 / MACH) ** 2 / GAMMA
MU, POWER, PR = 1.0 / 1600.0, 0.76, 0.72


def _sa():
    import spade_amd as sa
    sa.load()
    return sa


def _run(sa, make, steps, fused):
    """steps: advance() calls, each of 1 step, or (calls, steps per call)."""
    if fused:
        os.environ.pop("SPD_DISABLE_FUSED", None)
    else:
        os.environ["SPD_DISABLE_FUSED"] = "1"
    try:
        q, it = make(sa)
        assert it.fused == fused
        calls, per = steps
        for _ in range(calls):
            it.advance(per)
        out = q.download()
        it.close()
        return out
    finally:
        os.environ.pop("SPD_DISABLE_FUSED", None)


def _tgv(nb, nc, hot=False, stale_ghosts=False):
    def make(sa):
        g = sa.CartesianGrid(nc, nb, ((0.0, L),) * 3, num_exch=2)
        host = np.zeros(g.host_shape)
        if stale_ghosts:
            host[..., 0], host[..., 1] = R, 1.0       # uniform ghosts, not the exchanged values
        host[:, 2:-2, 2:-2, 2:-2] = bench.tgv_interior(nb, nc, L, R)
        if hot:
            # a warm patch in the first blocks: T / T_ref up to 1.4, outside the
            # power-law series range (|T/T_ref - 1| <= 1/16) at some faces
            inner = host[:, 2:-2, 2:-2, 2:-2]
            z = np.linspace(0.0, 1.0, nc[2])[None, :, None, None]
            inner[: max(1, inner.shape[0] // 3), ..., 1] *= 1.0 + 0.4 * z
        q = sa.GridArray(g).upload(host)
        if not stale_ghosts:
            sa.create_exchange(q).exchange(q)
        gas = sa.IdealGas(GAMMA, R)
        flux = sa.compose(sa.cent_keep(2, gas), sa.visc_lr(sa.power_law(MU, 1.0, POWER, PR), gas))
        assert sa.flux_div_kernel(q, flux) == "fast"
        dt = 0.5 * (L / (nb[0] * nc[0])) / sa.max_wavespeed(q, gas)
        return q, sa.Integrator(q, flux, dt)
    return make


@pytest.mark.parametrize("steps", [(1, 1), (2, 1), (1, 3)], ids=["1x1", "2x1", "1x3"])
def test_fused_tgv_equals_separate(steps):
    sa = _sa()
    make = _tgv((4, 4, 4), (16, 16, 16))
    print("fused vs separate: tgv 64^3, steps", steps, flush=True)
    a = _run(sa, make, steps, True)
    b = _run(sa, make, steps, False)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("nb,nc", [((1, 2, 3), (16, 16, 8)), ((2, 1, 2), (16, 16, 16))],
                         ids=["nz8_selfx", "selfy"])
def test_fused_block_shapes_equal_separate(nb, nc):
    sa = _sa()
    make = _tgv(nb, nc)
    print("fused vs separate: blocks", nb, nc, flush=True)
    assert np.array_equal(_run(sa, make, (2, 1), True), _run(sa, make, (2, 1), False))


def test_fused_reads_the_callers_ghosts_first():
    """The first stage of an advance() call reads the ghost cells the caller
    left (here a uniform state, not the exchanged values), later stages read
    the neighbour interiors, and the last stage writes the exchanged ghosts."""
    sa = _sa()
    make = _tgv((2, 2, 2), (16, 16, 16), stale_ghosts=True)
    a = _run(sa, make, (1, 2), True)
    b = _run(sa, make, (1, 2), False)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)


def test_fused_with_pow_fixup_equals_separate():
    sa = _sa()
    make = _tgv((4, 4, 4), (16, 16, 16), hot=True)
    print("fused vs separate: tgv 64^3 with out-of-range face temperatures", flush=True)
    a = _run(sa, make, (1, 2), True)
    b = _run(sa, make, (1, 2), False)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)


def test_fused_merged_fixup_equals_three_launches():
    """The fix-up of a fused stage in one launch (ghost pull, per-cell
    residual and stage per flagged block) equals its three separate launches
    (SPD_SPLIT_FIXUP=1), on a state with out-of-range face temperatures."""
    sa = _sa()
    make = _tgv((4, 4, 4), (16, 16, 16), hot=True)
    print("fused: merged vs three-launch fix-up", flush=True)
    a = _run(sa, make, (1, 3), True)
    os.environ["SPD_SPLIT_FIXUP"] = "1"
    try:
        b = _run(sa, make, (1, 3), True)
    finally:
        os.environ.pop("SPD_SPLIT_FIXUP", None)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)


@pytest.mark.parametrize("steps,hot", [((2, 1), False), ((1, 3), False), ((1, 2), True)],
                         ids=["2x1", "1x3", "1x2_fixup"])
def test_fused_channel_equals_separate(steps, hot):
    """The walled stretched channel: after the first stage of a call, the
    stages read their halo from the neighbour blocks and form the wall ghosts
    themselves; hot: blocks next to both walls whose face temperatures leave
    the power-law series range (their planes completed by the fix-up kernels,
    after the flagged blocks' ghosts are pulled and their walls filled)."""
    from oracle import pyoracle as po
    from spade_amd import studies
    sa = _sa()
    c = studies.channel()
    og = po.make_grid(c["nb"], c["nc"], c["ng"], c["lo"], c["hi"], periodic=(1, 0, 1),
                      maps=(None, (po.MAP_TANH,) + tuple(c["tanh_y"]), None))
    start = c["host"].copy()
    if hot:
        nbx, nby = c["nb"][0], c["nb"][1]
        for bx in range(3):
            start[bx + nbx * 0, 2:-2, 2:6, 2:-2, 1] *= 1.15                  # bottom row of blocks, near the wall
            start[bx + nbx * (nby - 1), 2:-2, -6:-2, 2:-2, 1] *= 1.15       # top row
    po.exchange(og, start)
    po.boundary_fill_wall(og, start, 2, c["T_wall"])
    po.boundary_fill_wall(og, start, 3, c["T_wall"])

    def make(sa):
        g = sa.CartesianGrid(c["nc"], c["nb"], tuple(zip(c["lo"], c["hi"])), num_exch=c["ng"],
                             periodic=(True, False, True), maps=(None, sa.tanh_map(*c["tanh_y"]), None))
        gas = sa.IdealGas(c["gamma"], c["R"])
        a, ep, et = c["pdiss"]
        flux = sa.compose(sa.cent_keep(2, gas), sa.pressure_diss_lr(gas, sa.const_sensor(a), ep, et),
                          sa.visc_lr(sa.power_law(c["mu"], 1.0, c["power"], c["prandtl"]), gas))
        q = sa.GridArray(g).upload(start)
        return q, sa.Integrator(q, flux, c["dt"], scheme=sa.ssprk3, walls=(2, 3), t_wall=c["T_wall"])
    print("fused vs separate: channel 192x128x96, steps", steps, "hot", hot, flush=True)
    a = _run(sa, make, steps, True)
    b = _run(sa, make, steps, False)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)


@pytest.mark.parametrize("steps", [(2, 1), (1, 3)], ids=["2x1", "1x3"])
def test_fused_periodic_stretched_equals_separate(steps):
    """A tanh-stretched grid periodic on every axis: the mapped kernel's
    stages read their halo from the neighbour blocks with no wall to form."""
    def make(sa):
        nb, nc = (2, 3, 2), (16, 16, 16)
        g = sa.CartesianGrid(nc, nb, ((0.0, L), (-1.0, 1.0), (0.0, L)), num_exch=2,
                             maps=(None, sa.tanh_map(-1.0, 1.0, 1.3), None))
        host = np.zeros(g.host_shape)
        host[:, 2:-2, 2:-2, 2:-2] = bench.tgv_interior(nb, nc, L, R)
        q = sa.GridArray(g).upload(host)
        sa.create_exchange(q).exchange(q)
        gas = sa.IdealGas(GAMMA, R)
        flux = sa.compose(sa.cent_keep(2, gas), sa.visc_lr(sa.power_law(MU, 1.0, POWER, PR), gas))
        assert sa.flux_div_kernel(q, flux) == "fast"
        return q, sa.Integrator(q, flux, 2e-4)
    sa = _sa()
    print("fused vs separate: periodic stretched grid, steps", steps, flush=True)
    a = _run(sa, make, steps, True)
    b = _run(sa, make, steps, False)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)


def test_not_fused_without_ghost_coverage():
    """A non-periodic axis without walls keeps the separate launches (its
    ghost cells are left to the caller, so both buffers would not agree)."""
    sa = _sa()
    g = sa.CartesianGrid((16, 16, 16), (2, 2, 2), ((0.0, 1.0),) * 3, num_exch=2, periodic=(True, False, True))
    q = sa.GridArray(g, fill=1.0)
    gas = sa.IdealGas(GAMMA, R)
    flux = sa.compose(sa.cent_keep(2, gas), sa.visc_lr(sa.power_law(MU, 1.0, POWER, PR), gas))
    assert not sa.Integrator(q, flux, 1e-4).fused
    assert sa.Integrator(q, flux, 1e-4, walls=(2, 3), t_wall=1.0).fused
    assert not sa.Integrator(q, flux, 1e-4, scheme=sa.ssprk3_opt, walls=(2, 3), t_wall=1.0).fused
