"""Kinetic energy and enstrophy against closed forms (the reference has no
such diagnostics, so nothing of it pins them).

The Taylor-Green vortex at t = 0 on [0, 2 pi]^3 with N_d cells along axis d,
u = u0 sin x cos y cos z, v = -u0 cos x sin y cos z, w = 0,
p = p0 + (rho0 u0^2 / 16)(cos 2x + cos 2y)(cos 2z + 2), T = T0, rho = p / (R T):

* kinetic energy <rho |u|^2 / 2> = rho0 u0^2 / 8.  With a = cos 2x and
  b = cos 2y, |u|^2 = u0^2 cos^2 z (1 - a b) / 2, and the pressure perturbation
  multiplies it by (a + b)(...), whose average over x and y vanishes:
  <(a + b)(1 - a b)> = 0.  Averages of these low modes over N > 4 equally
  spaced cell centres are exact, so the discrete value equals the continuous
  one up to the rounding of the sums.
* enstrophy <|omega|^2 / 2> with second-order central differences: the
  central difference of a wavenumber-one mode along axis d is the exact
  derivative times s_d = sin(h_d) / h_d, h_d = 2 pi / N_d; omega_x and
  omega_y each hold one z-derivative and omega_z = (s_x + s_y) sin x sin y
  cos z (times u0), so <|omega|^2 / 2> = u0^2 (2 s_z^2 + (s_x + s_y)^2) / 16
  (3/8 u0^2 s^2 when the axes agree).  Unequal N per axis make a swapped axis
  show.

Checked for spd_diagnostics (FP64 path; rho0 = u0 = T0 = 1, R = p0) and for
the first entry of the batched engine's per-step history (the state the first
step starts from; its TGV has rho0 = u0 = T0 = 1).
"""
import math

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu

L = 2.0 * np.pi
R = 100.0 / 1.4


def _closed(ncells):
    s = [math.sin(L / n) / (L / n) for n in ncells]
    return 0.125, (2.0 * s[2] ** 2 + (s[0] + s[1]) ** 2) / 16.0


def test_fp64_diagnostics_tgv_closed_form():
    import spade_amd as sa
    sa.load()
    nb, nc, ng = (2, 3, 2), (16, 16, 12), 2
    g = sa.CartesianGrid(nc, nb, ((0.0, L),) * 3, num_exch=ng)
    host = np.zeros(g.host_shape)
    host[:, ng:-ng, ng:-ng, ng:-ng] = bench.tgv_interior(nb, nc, L, R)
    q = sa.GridArray(g).upload(host)
    sa.create_exchange(q).exchange(q)
    ke, en = sa.diagnostics(q, sa.IdealGas(1.4, R))
    ke_ref, en_ref = _closed([nb[d] * nc[d] for d in range(3)])
    print("fp64 diagnostics", ke, ke_ref, en, en_ref, flush=True)
    assert abs(ke - ke_ref) <= 1e-13 * ke_ref
    assert abs(en - en_ref) <= 1e-13 * en_ref


def test_batched_diagnostics_tgv_closed_form():
    import spade_amd as sa
    sa.load()
    cells = (32, 48, 64)
    b = sa.Batch(256 << 20)
    cid = b.add(sa.CaseSpec("tgv", cells, mach=0.1, reynolds=1600.0))
    b.step(1)
    ke, en = b.diag(cid)
    ke_ref, en_ref = _closed(cells)
    print("batched diagnostics", ke[0], ke_ref, en[0], en_ref, flush=True)
    assert abs(ke[0] - ke_ref) <= 1e-13 * ke_ref
    assert abs(en[0] - en_ref) <= 1e-13 * en_ref


def test_max_wavespeed_equals_host_value():
    """max(|u| + sqrt(gamma R T)) over the interior cells (the CFL wave speed of
    development/cuda-tgv/main.cc:155-172): the same expression in the same
    order, correctly rounded square roots and an exact max, so the device value
    equals the host one bit for bit; twice, so the kept buffers are reused."""
    import spade_amd as sa
    sa.load()
    nb, nc, ng = (3, 2, 2), (16, 8, 12), 2
    g = sa.CartesianGrid(nc, nb, ((0.0, L),) * 3, num_exch=ng)
    rng = np.random.default_rng(11)
    gas = sa.IdealGas(1.4, R)
    for trial in range(2):
        host = np.zeros(g.host_shape)
        host[..., 0] = R * (1.0 + 0.1 * rng.random(host.shape[:-1]))
        host[..., 1] = 1.0 + 0.5 * rng.random(host.shape[:-1])
        host[..., 2:] = rng.standard_normal(host.shape[:-1] + (3,))
        q = sa.GridArray(g).upload(host)
        inner = host[:, ng:-ng, ng:-ng, ng:-ng]
        T, u, v, w = inner[..., 1], inner[..., 2], inner[..., 3], inner[..., 4]
        want = np.max(np.sqrt(1.4 * R * T) + np.sqrt(u * u + v * v + w * w))
        got = sa.max_wavespeed(q, gas)
        assert got == want, (trial, got, want)


def test_max_wavespeed_passes_nan_through():
    """A diverged state gives a NaN wave speed (the reference's driver stops on
    isnan(umax), development/cuda-tgv/main.cc:250-259): one NaN temperature in
    one interior cell of a block other than the first makes the reduction NaN."""
    import spade_amd as sa
    sa.load()
    g = sa.CartesianGrid((8, 8, 8), (2, 2, 2), ((0.0, 1.0),) * 3, num_exch=2)
    host = np.zeros(g.host_shape)
    host[..., 0] = 100.0
    host[..., 1] = 1.0
    host[..., 2] = 0.5
    q = sa.GridArray(g).upload(host)
    gas = sa.IdealGas(1.4, 100.0)
    assert np.isfinite(sa.max_wavespeed(q, gas))
    host[5, 4, 5, 6, 1] = np.nan           # block 5, interior cell (i, j, k) = (4, 3, 2)
    q.upload(host)
    assert np.isnan(sa.max_wavespeed(q, gas))
