"""The LDS-staged flux kernel with each block's planes split between two or
four workgroups (each with its own window prologue over [k0, k1), recording the
first plane it leaves to the fix-up in its own entry) equals the one-workgroup
launch bit for bit: overwrite and increment modes, face temperatures that
leave the power-law series range in the lower half, the upper half or both
(the fix-up then redoes the block from the first flagged plane, as the
one-workgroup launch does), the stretched channel (mapped coordinates), and
the fused SSP-RK3 step; two and four workgroups per block.  SPD_FAST_ZSPLIT=1|2|4
forces the split.  The workgroups taken in contiguous runs of blocks per XCD
(SPD_FAST_XCD=1, the default of the fused stages that read their halo from
the neighbour blocks) equal the launch order bit for bit.
"""
import os

import numpy as np
import pytest

import bench
from gpu_helpers import RawFlux

pytestmark = pytest.mark.gpu

L = 2.0 * np.pi
GAMMA, R = 1.4, (1.0 / 0.1) ** 2 / 1.4
K2 = dict(kind=1, order=2)
VP = dict(kind=7, law=1, mu=1.0 / 1600.0, T_ref=1.0, power=0.76, prandtl=0.72)


def _sa():
    import spade_amd as sa
    sa.load()
    return sa


def _same(a, b):
    """bit-equal; otherwise print where (block, plane) the arrays differ"""
    if np.array_equal(a, b):
        return True
    d = np.argwhere(a != b)
    rel = np.max(np.abs(a - b)) / np.max(np.abs(b))
    print("differing entries", len(d), "max rel", rel, "blocks", np.unique(d[:, 0])[:20],
          "planes", np.unique(d[:, 1]), "vars", np.unique(d[:, 4]), flush=True)
    return False


def _with_split(zs, fn):
    os.environ["SPD_FAST_ZSPLIT"] = str(zs)
    try:
        return fn()
    finally:
        os.environ.pop("SPD_FAST_ZSPLIT", None)


def _tgv(sa, nb=(4, 4, 4), hot_planes=()):
    nc = (16, 16, 16)
    g = sa.CartesianGrid(nc, nb, ((0.0, L),) * 3, num_exch=2)
    host = np.zeros(g.host_shape)
    host[:, 2:-2, 2:-2, 2:-2] = bench.tgv_interior(nb, nc, L, R)
    for blk, k in hot_planes:            # T / T_ref = 1.3 on one plane of one block
        host[blk, 2 + k, 2:-2, 2:-2, 1] *= 1.3
    q = sa.GridArray(g).upload(host)
    sa.create_exchange(q).exchange(q)
    return g, q


@pytest.mark.parametrize("hot", [(), ((0, 3),), ((1, 12),), ((2, 5), (2, 13)), ((3, 7),), ((3, 8),)],
                         ids=["none", "lower", "upper", "both", "at7", "at8"])
@pytest.mark.parametrize("increment", [False, True])
@pytest.mark.parametrize("zs", [2, 4])
def test_split_flux_div_equals_whole(hot, increment, zs):
    sa = _sa()
    g, q = _tgv(sa, hot_planes=hot)
    init = np.random.default_rng(5).standard_normal(g.host_shape)
    print("zsplit", zs, "flux_div, hot planes", hot, "increment", increment, flush=True)

    def run():
        rhs = sa.GridArray(g).upload(init)
        assert sa.flux_div_kernel(q, RawFlux([K2, VP], GAMMA, R)) == "fast"
        sa.flux_div(q, rhs, RawFlux([K2, VP], GAMMA, R), increment=increment)
        return rhs.download()
    a, b = _with_split(1, run), _with_split(zs, run)
    assert np.isfinite(a).all()
    assert _same(a, b)


@pytest.mark.parametrize("zs", [2, 4])
def test_split_channel_flux_div_equals_whole(zs):
    from spade_amd import studies
    sa = _sa()
    c = studies.channel()
    g = sa.CartesianGrid(c["nc"], c["nb"], tuple(zip(c["lo"], c["hi"])), num_exch=c["ng"],
                         periodic=(True, False, True), maps=(None, sa.tanh_map(*c["tanh_y"]), None))
    q = sa.GridArray(g).upload(c["host"])
    sa.create_exchange(q).exchange(q)
    sa.boundary_fill_wall(q, 2, c["T_wall"])
    sa.boundary_fill_wall(q, 3, c["T_wall"])
    gas = sa.IdealGas(c["gamma"], c["R"])
    a0, ep, et = c["pdiss"]
    flux = sa.compose(sa.cent_keep(2, gas), sa.pressure_diss_lr(gas, sa.const_sensor(a0), ep, et),
                      sa.visc_lr(sa.power_law(c["mu"], 1.0, c["power"], c["prandtl"]), gas))
    print("zsplit flux_div, channel 192x128x96", flush=True)

    def run():
        rhs = sa.GridArray(g, fill=0.0)
        assert sa.flux_div_kernel(q, flux) == "fast"
        sa.flux_div(q, rhs, flux, increment=False)
        return rhs.download()
    a, b = _with_split(1, run), _with_split(zs, run)
    assert np.isfinite(a).all()
    assert _same(a, b)


@pytest.mark.parametrize("zs", [2, 4])
@pytest.mark.parametrize("hot", [(), ((1, 9),)], ids=["plain", "fixup"])
def test_split_fused_step_equals_whole(hot, zs):
    sa = _sa()
    print("zsplit fused step, hot planes", hot, flush=True)

    def run():
        g, q = _tgv(sa, nb=(2, 2, 2), hot_planes=hot)
        gas = sa.IdealGas(GAMMA, R)
        flux = sa.compose(sa.cent_keep(2, gas), sa.visc_lr(sa.power_law(1.0 / 1600.0, 1.0, 0.76, 0.72), gas))
        it = sa.Integrator(q, flux, 1e-3)
        assert it.fused
        it.advance(2)
        out = q.download()
        it.close()
        return out
    a, b = _with_split(1, run), _with_split(zs, run)
    assert np.isfinite(a).all()
    assert _same(a, b)


def _with_xcd(x, fn):
    os.environ["SPD_FAST_XCD"] = str(x)
    try:
        return fn()
    finally:
        os.environ.pop("SPD_FAST_XCD", None)


@pytest.mark.parametrize("zs", [1, 4])
def test_xcd_runs_flux_div_equals_launch_order(zs):
    """Workgroups taken in contiguous runs of blocks per XCD (a permutation of
    the launch order) give the same residual as the launch order, with and
    without a block split and with a block left to the fix-up; the residual
    array starts as noise, so a block that no workgroup took would show."""
    sa = _sa()
    g, q = _tgv(sa, nb=(4, 3, 5), hot_planes=((7, 9),))
    init = np.random.default_rng(11).standard_normal(g.host_shape)
    print("xcd runs flux_div, zsplit", zs, flush=True)

    def run():
        rhs = sa.GridArray(g).upload(init)
        sa.flux_div(q, rhs, RawFlux([K2, VP], GAMMA, R), increment=False)
        return rhs.download()
    a = _with_split(zs, lambda: _with_xcd(0, run))
    b = _with_split(zs, lambda: _with_xcd(1, run))
    assert np.isfinite(a).all()
    assert _same(a, b)


@pytest.mark.parametrize("hot", [(), ((1, 9),)], ids=["plain", "fixup"])
def test_xcd_runs_fused_step_equals_launch_order(hot):
    """The fused SSP-RK3 stages (whose neighbour-reading launches take XCD
    runs by default) equal the same stages in launch order, bit for bit."""
    sa = _sa()
    print("xcd runs fused step, hot planes", hot, flush=True)

    def run():
        g, q = _tgv(sa, nb=(3, 2, 2), hot_planes=hot)
        gas = sa.IdealGas(GAMMA, R)
        flux = sa.compose(sa.cent_keep(2, gas), sa.visc_lr(sa.power_law(1.0 / 1600.0, 1.0, 0.76, 0.72), gas))
        it = sa.Integrator(q, flux, 1e-3)
        assert it.fused
        it.advance(3)
        out = q.download()
        it.close()
        return out
    a, b = _with_xcd(0, run), _with_xcd(1, run)
    assert np.isfinite(a).all()
    assert _same(a, b)
