"""spade::io on device arrays, through the C ABI:

* io::output_vtk of the golden TGV state uploaded to the GPU writes the same
  bytes as the compiled reference's io/io_vtk.h (tests/golden/vtk_g1);
* io::binary_write writes the array's host image (reference layout, ghosts
  included, io/io_native.h:19-45): on G1 the file equals g1_prim_ghost.bin;
* binary_read(binary_write(a)) restores a into a fresh array bit for bit, on a
  grid large enough to go through the staging buffer in several chunks, and
  output_vtk of that device array equals output_vtk_host of its download;
* a file of the wrong size is refused with SPD_ERR_IO.
"""
import os

import numpy as np
import pytest

import golden_io as gi
from cases import G1

pytestmark = pytest.mark.gpu


def _sa():
    import spade_amd as sa
    sa.load()
    return sa


def _g1(sa):
    return sa.CartesianGrid(G1["nc"], G1["nb"], ((0.0, gi.TWO_PI_REF),) * 3, num_exch=G1["ng"])


def _same_tree(a, b):
    fa = sorted(os.path.relpath(os.path.join(d, f), a) for d, _, fs in os.walk(a) for f in fs)
    fb = sorted(os.path.relpath(os.path.join(d, f), b) for d, _, fs in os.walk(b) for f in fs)
    assert fa == fb
    for rel in fa:
        with open(os.path.join(a, rel), "rb") as x, open(os.path.join(b, rel), "rb") as y:
            assert x.read() == y.read(), rel


def test_output_vtk_matches_reference_bytes(tmp_path):
    sa = _sa()
    q = sa.GridArray(_g1(sa)).upload(gi.load("g1_prim_ghost"))
    sa.output_vtk(str(tmp_path), "g1", q)
    _same_tree(os.path.join(gi.GOLDEN, "vtk_g1"), str(tmp_path))


def test_binary_write_is_the_host_image(tmp_path):
    sa = _sa()
    host = gi.load("g1_prim_ghost")
    q = sa.GridArray(_g1(sa)).upload(host)
    path = tmp_path / "g1.bin"
    sa.binary_write(str(path), q)
    assert path.read_bytes() == np.ascontiguousarray(host, dtype="<f8").tobytes()


def test_binary_round_trip_and_vtk_in_chunks(tmp_path):
    sa = _sa()
    # 10 x 10 x 12 blocks of 8^3 (12^3 stored): 82.9 MB, more than one 64 MiB staging chunk
    g = sa.CartesianGrid((8, 8, 8), (10, 10, 12), ((0.0, 1.0), (0.0, 2.0), (0.0, 3.0)), num_exch=2)
    rng = np.random.default_rng(7)
    host = rng.standard_normal(g.host_shape)
    print("io: grid", g.num_blocks, "image", host.nbytes, "bytes", flush=True)
    a = sa.GridArray(g).upload(host)
    path = tmp_path / "a.bin"
    sa.binary_write(str(path), a)
    assert os.path.getsize(path) == host.nbytes
    b = sa.GridArray(g, fill=0.0)
    sa.binary_read(str(path), b)
    assert np.array_equal(b.download(), host)
    sa.output_vtk(str(tmp_path / "dev"), "a", b, names=sa.CONS_NAMES)
    sa.output_vtk_host(str(tmp_path / "host"), "a", g.desc, host, names=sa.CONS_NAMES)
    _same_tree(str(tmp_path / "dev"), str(tmp_path / "host"))
    print("io: round trip and", g.nblocks, "block files equal", flush=True)


def test_binary_read_rejects_wrong_size(tmp_path):
    sa = _sa()
    q = sa.GridArray(_g1(sa), fill=1.0)
    path = tmp_path / "short.bin"
    path.write_bytes(b"\0" * 64)
    with pytest.raises(sa.SpadeError) as e:
        sa.binary_read(str(path), q)
    assert e.value.code == 5
    assert np.all(q.download() == 1.0)
