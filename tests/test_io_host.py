"""spade::io formats written by the library's host-side writer (no device):

* io::output_vtk of the golden TGV state on grid G1 equals, byte for byte,
  the files the compiled reference wrote for the same array (tests/golden/
  vtk_g1, made by make_golden.sh through io/io_vtk.h:276-288);
* independently of those bytes: every DataArray payload decodes (Python's
  base64) to a 4-byte count and the data the reference layout fixes (interior
  cells of one variable, i fastest; node coordinates box min + i dx);
* node coordinates of a tanh-stretched axis: the map is the identity at the
  ends of the stretched interval, increasing, and its centred differences
  match the metric derivative beta (1 - t^2)/tanh(beta) the flux kernels use;
* error behaviour: unwritable paths give SPD_ERR_IO, bad names ValueError.
"""
import base64
import math
import os
import re

import numpy as np
import pytest

import golden_io as gi
from cases import G1

GOLD_VTK = os.path.join(gi.GOLDEN, "vtk_g1")


def _sa():
    import spade_amd as sa
    sa.load()
    return sa


def _g1_desc(sa, maps=(None, None, None)):
    return sa.grid_desc(G1["nc"], G1["nb"], ((0.0, gi.TWO_PI_REF),) * 3, num_exch=G1["ng"], maps=maps)


def _files(root):
    out = {}
    for d, _, fs in os.walk(root):
        for f in fs:
            p = os.path.join(d, f)
            out[os.path.relpath(p, root)] = p
    return out


def _payloads(text):
    """(name, decoded count, decoded float64 data) of every DataArray."""
    out = []
    for m in re.finditer(r'<DataArray type="Float64" (?:Name="([^"]*)" )?(?:NumberOfComponents="3" )?'
                         r'format="binary">\n([^\n]*)\n</DataArray>', text):
        name, b64 = m.group(1), m.group(2)
        count = int.from_bytes(base64.b64decode(b64[:8]), "little")
        data = np.frombuffer(base64.b64decode(b64[8:]), dtype="<f8")
        out.append((name, count, data))
    return out


def test_vtk_matches_reference_bytes(tmp_path):
    sa = _sa()
    sa.output_vtk_host(str(tmp_path), "g1", _g1_desc(sa), gi.load("g1_prim_ghost"))
    want, got = _files(GOLD_VTK), _files(str(tmp_path))
    assert sorted(want) == sorted(got)
    for rel, p in want.items():
        with open(p, "rb") as a, open(got[rel], "rb") as b:
            assert a.read() == b.read(), rel


def test_vtk_payloads_decode_to_the_array(tmp_path):
    sa = _sa()
    host = gi.load("g1_prim_ghost")
    sa.output_vtk_host(str(tmp_path), "g1", _g1_desc(sa), host, names=sa.FLUX_NAMES)
    ng = G1["ng"]
    nx, ny, nz = G1["nc"]
    inner = host[:, ng:-ng, ng:-ng, ng:-ng, :]
    nbx, nby, _ = G1["nb"]
    h = gi.TWO_PI_REF
    base = open(os.path.join(tmp_path, "g1.pvts")).read()
    assert 'Scalars = "continuity,energy,x_momentum,y_momentum,z_momentum"' in base
    assert f'WholeExtent="0 {nx} 0 {ny} 0 {nz * inner.shape[0]}"' in base
    for lb in range(inner.shape[0]):
        text = open(os.path.join(tmp_path, "data_g1", f"b{lb:09d}.vts")).read()
        arrays = _payloads(text)
        assert [a[0] for a in arrays] == list(sa.FLUX_NAMES) + [None]
        for v, (_, count, data) in enumerate(arrays[:5]):
            assert count == 8 * nx * ny * nz
            assert np.array_equal(data, inner[lb, ..., v].reshape(-1))
        _, count, pts = arrays[5]
        assert count == 8 * 3 * (nx + 1) * (ny + 1) * (nz + 1)
        pts = pts.reshape(nz + 1, ny + 1, nx + 1, 3)
        b = (lb % nbx, (lb // nbx) % nby, lb // (nbx * nby))
        for a, n in enumerate((nx, ny, nz)):
            bsize = h / G1["nb"][a]
            mn = 0.0 + b[a] * bsize
            dx = ((mn + bsize) - mn) / n
            idx = np.arange(n + 1, dtype=np.float64)
            want = mn + idx * dx
            got = np.moveaxis(pts[..., a], 2 - a, 0)
            assert np.array_equal(got, np.broadcast_to(want.reshape(-1, 1, 1), got.shape)), (lb, a)


def test_vtk_stretched_axis_nodes(tmp_path):
    sa = _sa()
    beta = 1.5
    desc = _g1_desc(sa, maps=(None, sa.tanh_map(0.0, gi.TWO_PI_REF, beta), None))
    host = gi.load("g1_prim_ghost")
    sa.output_vtk_host(str(tmp_path), "s", desc, host)
    nx, ny, nz = G1["nc"]
    nbx, nby, _ = G1["nb"]
    ys = {}
    for lb in range(host.shape[0]):
        text = open(os.path.join(tmp_path, "data_s", f"b{lb:09d}.vts")).read()
        pts = _payloads(text)[5][2].reshape(nz + 1, ny + 1, nx + 1, 3)
        by = (lb // nbx) % nby
        ys[by] = pts[0, :, 0, 1]
    y = np.concatenate([ys[b][:-1] for b in range(nby)] + [ys[nby - 1][-1:]])
    assert abs(y[0]) <= 1e-15 * gi.TWO_PI_REF and abs(y[-1] - gi.TWO_PI_REF) <= 1e-15 * gi.TWO_PI_REF
    assert np.all(np.diff(y) > 0)
    # centred differences against the metric x'(xi) at the cell centres
    c, hw = 0.5 * gi.TWO_PI_REF, 0.5 * gi.TWO_PI_REF
    dxi = gi.TWO_PI_REF / (nby * ny)
    xi_mid = (np.arange(len(y) - 1) + 0.5) * dxi
    t = np.tanh(beta * (xi_mid - c) / hw)
    deriv = beta * (1 - t * t) / math.tanh(beta)
    assert np.allclose(np.diff(y) / dxi, deriv, rtol=1e-2)


def test_vtk_errors(tmp_path):
    sa = _sa()
    host = gi.load("g1_prim_ghost")
    blocker = tmp_path / "file"
    blocker.write_text("not a directory")
    with pytest.raises(sa.SpadeError) as e:
        sa.output_vtk_host(str(blocker), "g1", _g1_desc(sa), host)
    assert e.value.code == 5
    with pytest.raises(ValueError):
        sa.output_vtk_host(str(tmp_path), "g1", _g1_desc(sa), host, names=("a", "b"))
    with pytest.raises(ValueError):
        sa.output_vtk_host(str(tmp_path), "g1", _g1_desc(sa), host[:-1])
