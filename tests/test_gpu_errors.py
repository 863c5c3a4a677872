"""Error behaviour of the C ABI: a non-zero status with spd_last_error(), raised
as SpadeError by the Python mirror (the reference throws except::sp_exception
for bad sizes and configurations, core/except.h:12-30).  Every case here is
refused on the host before any kernel is launched; the arrays it was given are
left as they were, and the library keeps working afterwards.
"""
import numpy as np
import pytest

from gpu_helpers import RawFlux

pytestmark = pytest.mark.gpu

GAMMA, R = 1.4, 100.0 / 1.4
ARG = 1
K2 = dict(kind=1, order=2)
VC = dict(kind=7, law=0, mu=1e-3, prandtl=0.72)


def _sa():
    import spade_amd as sa
    sa.load()
    return sa


def _arrays(sa, ng=2, nc=(8, 8, 8)):
    g = sa.CartesianGrid(nc, (2, 1, 1), ((0.0, 1.0),) * 3, num_exch=ng)
    rng = np.random.default_rng(3)
    host = np.zeros(g.host_shape)
    host[..., 0] = R * (1 + 0.01 * rng.standard_normal(host.shape[:-1]))
    host[..., 1] = 1.0
    host[..., 2:] = 0.1 * rng.standard_normal(host.shape[:-1] + (3,))
    q = sa.GridArray(g).upload(host)
    rhs = sa.GridArray(g, fill=7.0)
    return g, q, rhs


def _refused(sa, fn, words):
    with pytest.raises(sa.SpadeError) as e:
        fn()
    assert e.value.code == ARG, e.value
    assert words in str(e.value), str(e.value)


def test_grid_blocks_narrower_than_ghost_layers():
    sa = _sa()
    _refused(sa, lambda: sa.CartesianGrid((8, 8, 1), (1, 1, 2), ((0.0, 1.0),) * 3, num_exch=2),
             "at least num_exch")


@pytest.mark.parametrize("term, words", [
    (dict(kind=10, order=2), "unknown flux term kind"),
    (dict(kind=8, order=2), "hybrid dissipative scheme"),                    # diss_kind missing
    (dict(kind=8, order=6, diss_kind=4), "hybrid central scheme"),
    (dict(kind=1, order=6), "order/2 ghost layers"),
    (dict(kind=1, order=3), "even"),
    (dict(kind=7, law=2, mu=1e-3, prandtl=0.72), "viscous law"),
    (dict(kind=7, law=0, mu=1e-3, prandtl=0.0), "Prandtl"),
    (dict(kind=7, law=1, mu=1e-3, T_ref=0.0, power=0.7, prandtl=0.72), "T_ref"),
    (dict(kind=7, law=0, mu=1e-3, prandtl=0.72, sgs=2), "subgrid model"),
    (dict(kind=7, law=0, mu=1e-3, prandtl=0.72, sgs=1, sgs_cw=0.5, sgs_delta=0.1, sgs_prt=0.0), "turbulent Prandtl"),
    (dict(kind=3, sensor=5, sensor_param=0.1, eps_p=0.1, eps_T=0.1), "sensor"),
    (dict(kind=4, split=3, smooth=1), "split"),
])
def test_flux_descriptor_refused(term, words):
    sa = _sa()
    g, q, rhs = _arrays(sa)
    _refused(sa, lambda: sa.flux_div(q, rhs, RawFlux([term], GAMMA, R), increment=False), words)
    assert np.all(rhs.download() == 7.0)
    # still working: a valid call on the same arrays
    sa.flux_div(q, rhs, RawFlux([K2, VC], GAMMA, R), increment=False)
    assert np.all(np.isfinite(rhs.download()))


def test_weno_needs_two_ghost_layers():
    sa = _sa()
    g, q, rhs = _arrays(sa, ng=1)
    for t in (dict(kind=4, split=0, smooth=1), dict(kind=5, smooth=1), dict(kind=9, smooth=0)):
        _refused(sa, lambda: sa.flux_div(q, rhs, RawFlux([t], GAMMA, R)), "2 ghost layers")


def test_arrays_of_different_grids_or_the_same_array():
    sa = _sa()
    g, q, rhs = _arrays(sa)
    g2, q2, rhs2 = _arrays(sa)
    flux = RawFlux([K2], GAMMA, R)
    _refused(sa, lambda: sa.flux_div(q, rhs2, flux), "different grids")
    _refused(sa, lambda: sa.flux_div(q, q, flux), "different arrays")
    assert np.all(rhs2.download() == 7.0)


def test_wall_face_out_of_range():
    sa = _sa()
    g, q, rhs = _arrays(sa)
    before = q.download()
    _refused(sa, lambda: sa.boundary_fill_wall(q, 6, 1.0), "face")
    assert np.array_equal(q.download(), before)


def test_hybrid_refused_on_the_device():
    """A well-formed hybrid_scheme_t descriptor, alone or composed, is a valid
    request the device library does not implement: SPD_ERR_UNSUPPORTED, before
    any launch, arrays untouched."""
    sa = _sa()
    g, q, rhs = _arrays(sa)
    hyb = dict(kind=8, order=2, diss_kind=4, split=0, smooth=1, sensor=0, sensor_param=0.5, full=1)
    for terms in ([hyb], [hyb, VC]):
        with pytest.raises(sa.SpadeError) as e:
            sa.flux_div(q, rhs, RawFlux(terms, GAMMA, R), increment=False)
        assert e.value.code == 3 and "hybrid" in str(e.value), str(e.value)
    assert np.all(rhs.download() == 7.0)


def test_boundary_fill_arguments_refused():
    sa = _sa()
    g, q, rhs = _arrays(sa)               # 8^3 cells per block
    before = q.download()
    _refused(sa, lambda: sa.boundary_fill(q, 64, sa.extrapolate(1)), "face mask")
    _refused(sa, lambda: sa.boundary_fill(q, sa.xmin, sa.extrapolate(8)), "order + 1")
    _refused(sa, lambda: sa.boundary_fill(q, sa.xmin, sa.extrapolate(-1)), "order")
    assert np.array_equal(q.download(), before)
