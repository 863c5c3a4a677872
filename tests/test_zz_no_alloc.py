"""spd_max_wavespeed_noalloc: the CFL reduction with its per-block partials in
the library's static device scratch and the result left in the caller's device
memory (no allocation).  Called through the C ABI and checked against
spd_max_wavespeed on two calls in a row, the second on a smaller grid, bit for
bit (a maximum has no rounding); an oversized grid and a host pointer are
refused before any launch.

The caller's device memory comes from the HIP runtime the library itself
uses (libamdhip64 from the ROCm install): torch ships a runtime of its own,
and two runtimes in one process do not share the device."""
import ctypes as C

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu


def _sa():
    import spade_amd as sa
    sa.load()
    return sa


def _tgv(sa, nb, nc, R):
    L = 2.0 * np.pi
    g = sa.CartesianGrid(nc, nb, ((0.0, L),) * 3, num_exch=2)
    host = np.zeros(g.host_shape)
    host[:, 2:-2, 2:-2, 2:-2] = bench.tgv_interior(nb, nc, L, R)
    host[:, 2:-2, 2:-2, 2:-2, 1] += 0.01 * np.sin(np.arange(host[:, 2:-2, 2:-2, 2:-2, 1].size)).reshape(
        host[:, 2:-2, 2:-2, 2:-2, 1].shape)
    q = sa.GridArray(g).upload(host)
    return g, q


class DeviceBuffer:
    """Device memory of the caller, from the library's own HIP runtime."""

    def __init__(self, ndouble):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(8 * ndouble)) == 0
        self.n = ndouble
        self.write(np.full(ndouble, -1.0))

    def write(self, host):
        h = np.ascontiguousarray(host, dtype=np.float64)
        assert self.hip.hipMemcpy(self.ptr, h.ctypes.data_as(C.c_void_p), C.c_size_t(8 * self.n), 1) == 0

    def read(self):
        h = np.empty(self.n)
        assert self.hip.hipMemcpy(h.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(8 * self.n), 2) == 0
        return h

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def test_noalloc_equals_allocating_reduction_twice():
    sa = _sa()
    R = 100.0 / 1.4
    gas = sa.IdealGas(1.4, R)
    sa.default_context()
    out = DeviceBuffer(2)
    try:
        for n, (nb, nc) in enumerate((((4, 4, 4), (16, 16, 16)), ((2, 2, 2), (8, 8, 8)))):
            g, q = _tgv(sa, nb, nc, R)
            want = sa.max_wavespeed(q, gas)
            print("case", n, "blocks", nb, "cells", nc, "reference", want, flush=True)
            sa.max_wavespeed_noalloc(q, gas, out.ptr.value + 8 * n)
            g.ctx.sync()
            got = out.read()[n]
            assert got == want, (n, got, want)
    finally:
        out.free()


def test_noalloc_refuses_oversized_grid_and_host_pointer():
    sa = _sa()
    gas = sa.IdealGas(1.4, 100.0)
    sa.default_context()
    out = DeviceBuffer(1)
    try:
        out.write(np.zeros(1))
        big = sa.CartesianGrid((2, 2, 2), (33, 33, 31), ((0.0, 1.0),) * 3, num_exch=1)   # 33759 blocks > 32768
        q = sa.GridArray(big, fill=1.0)
        with pytest.raises(sa.SpadeError) as e:
            sa.max_wavespeed_noalloc(q, gas, out.ptr.value)
        assert "at most 32768 blocks" in str(e.value)
        g, q2 = _tgv(sa, (2, 2, 2), (8, 8, 8), 100.0 / 1.4)
        host = np.zeros(1)
        with pytest.raises(sa.SpadeError) as e:
            sa.max_wavespeed_noalloc(q2, gas, host.ctypes.data)
        assert "device memory" in str(e.value)
        g.ctx.sync()
        assert out.read()[0] == 0.0
        # the library keeps working after the refusals
        assert sa.max_wavespeed(q2, gas) > 0.0
    finally:
        out.free()
