"""Readers for the golden fixtures in tests/golden (made by make_golden.sh).

Fixture files are raw little-endian float64 in the SPADE reference order
(block, k, j, i, var); names ending in "_ghost" hold the ghost layers too.
The domain of every fixture grid is [0, 2*3.1415926535]^3: the reference's pi
constant is truncated to ten decimals (core/consts.h:7).
"""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TWO_PI_REF = 2.0 * 3.1415926535


def manifest():
    files, params = {}, {}
    with open(os.path.join(GOLDEN, "manifest.txt")) as fh:
        for line in fh:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "file":
                files[tok[1]] = tuple(int(t) for t in tok[2:])
            elif tok[0] == "param":
                params[tok[1]] = float(tok[2])
    return files, params


def load(name):
    files, _ = manifest()
    shape = files[name]
    return np.fromfile(os.path.join(GOLDEN, name + ".bin"), dtype="<f8").reshape(shape)


def embed(interior, ng):
    """Interior-only fixture -> ghosted array (ghosts zero)."""
    nb, nz, ny, nx, nv = interior.shape
    out = np.zeros((nb, nz + 2 * ng, ny + 2 * ng, nx + 2 * ng, nv))
    out[:, ng:ng + nz, ng:ng + ny, ng:ng + nx, :] = interior
    return out


def interior(a, ng):
    return a[:, ng:-ng, ng:-ng, ng:-ng, :]


def rel_maxnorm(a, b):
    """max |a-b| / max |b| per variable (last axis)."""
    a = a.reshape(-1, a.shape[-1])
    b = b.reshape(-1, b.shape[-1])
    scale = np.max(np.abs(b), axis=0)
    scale[scale == 0] = 1.0
    return np.max(np.abs(a - b), axis=0) / scale
