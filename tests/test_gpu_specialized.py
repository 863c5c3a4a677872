"""The per-cell kernel specialised at compile time for a composition runs the
same arithmetic in the same order as the generic per-cell kernel, so on the
same input the two give the same residual bit for bit.  Compositions: every
non-empty subsequence of cent_keep<2>, pressure_diss_lr and visc_lr (in that
order), and cent_keep<4>, totani_lr, weno_t (Rusanov and physical splitting,
smooth and linear weights), fweno_t, first_order_t and weno_fds_t (linear
weights), each alone and followed
by visc_lr.  Checked on the stretched wall-bounded grid (where the LDS-staged kernel
does not apply) and on an identity grid with 8^3 blocks (nor there), in
overwrite and increment modes; SPD_DISABLE_SPEC=1 selects the generic kernel.
Parity with the oracle for these compositions on stretched grids is covered
by test_gpu_curvilinear.py and test_gpu_channel.py, which now run this kernel.
"""
import itertools
import os

import numpy as np
import pytest

from gpu_helpers import RawFlux
from test_gpu_curvilinear import TANH_Y, _grids, _sa, _state
from cases import GAMMA, R

pytestmark = pytest.mark.gpu

KEEP = dict(kind=1, order=2)
PDISS = [dict(kind=3, sensor=0, sensor_param=0.3, eps_p=0.1, eps_T=0.05),
         dict(kind=3, sensor=1, sensor_param=1e-6, eps_p=0.1, eps_T=0.05)]
VISC = [dict(kind=7, law=0, mu=2e-3, prandtl=0.72),
        dict(kind=7, law=1, mu=2e-3, T_ref=1.1, power=0.76, prandtl=0.72)]


OTHERS = [("keep4", dict(kind=1, order=4)), ("totani", dict(kind=2)),
          ("weno_rus", dict(kind=4, split=0, smooth=1)), ("weno_rus_lin", dict(kind=4, split=0, smooth=0)),
          ("weno_phys", dict(kind=4, split=1, smooth=1)), ("fweno", dict(kind=5, smooth=1)),
          ("fweno_lin", dict(kind=5, smooth=0)), ("first_order", dict(kind=6)),
          ("weno_fds_lin", dict(kind=9, smooth=0))]


def _compositions():
    out = []
    for name, t in OTHERS:
        out.append((name, [t]))
        out.append((name + "+visc1", [t, VISC[1]]))
    for k, p, v in itertools.product((None, KEEP), [None] + PDISS, [None] + VISC):
        terms = [t for t in (k, p, v) if t is not None]
        if terms:
            name = "+".join(["keep2"] * (k is not None)
                            + ([f"pdiss{p['sensor']}"] if p else []) + ([f"visc{v['law']}"] if v else []))
            out.append((name, terms))
    return out


def _run(sa, q, terms, increment, generic):
    rhs = sa.GridArray(q.grid, fill=0.25)
    if generic:
        os.environ["SPD_DISABLE_SPEC"] = "1"
    try:
        kind = sa.flux_div_kernel(q, RawFlux(terms, GAMMA, R))
        sa.flux_div(q, rhs, RawFlux(terms, GAMMA, R), increment=increment)
    finally:
        os.environ.pop("SPD_DISABLE_SPEC", None)
    return kind, rhs.download()


@pytest.mark.parametrize("geometry", ["stretched", "identity8"])
@pytest.mark.parametrize("name, terms", _compositions(), ids=[c[0] for c in _compositions()])
def test_specialized_equals_generic(geometry, name, terms):
    sa = _sa()
    dg, og = _grids(sa, maps=(None, TANH_Y, None) if geometry == "stretched" else (None, None, None))
    q = sa.GridArray(dg).upload(_state(og))
    for increment in (False, True):
        ks, spec = _run(sa, q, terms, increment, generic=False)
        kg, gen = _run(sa, q, terms, increment, generic=True)
        print(geometry, name, increment, ks, kg, flush=True)
        assert (ks, kg) == ("specialized", "generic")
        assert np.all(np.isfinite(spec))
        assert np.array_equal(spec, gen), np.max(np.abs(spec - gen))


@pytest.mark.parametrize("increment", [False, True], ids=["overwrite", "increment"])
def test_fixup_specialized_equals_generic(increment):
    """The fix-up launch after the LDS-staged kernel (blocks whose face
    temperatures leave the power law's series range) uses the specialised
    per-cell code too: same bits as with the generic per-cell code."""
    from test_gpu_fast import _state as fast_state
    sa = _sa()
    nb, nc, ng = (2, 2, 2), (16, 16, 16), 2
    host, L = fast_state(nb, nc, ng, 11)
    inner = host[:, ng:-ng, ng:-ng, ng:-ng]
    inner[..., 1] = 1.0 + 0.005 * np.random.default_rng(3).standard_normal(inner.shape[:-1])
    inner[1, 5, 3, 4, 1] = 1.4
    inner[5, :, :, :, 1] = 1.25
    g = sa.CartesianGrid(nc, nb, ((0.0, L),) * 3, num_exch=ng)
    q = sa.GridArray(g).upload(host)
    sa.create_exchange(q).exchange(q)
    terms = [dict(kind=1, order=2), VISC[1]]
    init = np.random.default_rng(5).standard_normal(g.host_shape)
    out = []
    for generic in (False, True):
        rhs = sa.GridArray(g).upload(init)
        if generic:
            os.environ["SPD_DISABLE_SPEC"] = "1"
        try:
            assert sa.flux_div_kernel(q, RawFlux(terms, GAMMA, R)) == "fast"
            sa.flux_div(q, rhs, RawFlux(terms, GAMMA, R), increment=increment)
        finally:
            os.environ.pop("SPD_DISABLE_SPEC", None)
        out.append(rhs.download())
    print("fixup", increment, np.max(np.abs(out[0] - out[1])), flush=True)
    assert np.all(np.isfinite(out[0]))
    assert np.array_equal(out[0], out[1])
