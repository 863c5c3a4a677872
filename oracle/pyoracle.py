"""ctypes binding of the CPU oracle (test infrastructure only).

Only the tests, the smoke check and the CPU-baseline leg of bench.py use this
module, and only as the checker.  Arrays are numpy float64 in the SPADE
reference layout, shape (nblocks, Nz, Ny, Nx, 5) with ghosts stored
(core/mem_map.h:466-497, 654-676).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

MAP_IDENTITY, MAP_TANH = 0, 1
TERM_KEEP, TERM_TOTANI, TERM_PDISS, TERM_WENO, TERM_FWENO, TERM_FIRST, TERM_VISC, TERM_HYBRID, TERM_WENO_FDS = \
    range(1, 10)
SENSOR_CONST, SENSOR_DUCROS = 0, 1
SPLIT_RUSANOV, SPLIT_PHYS = 0, 1
LAW_CONST, LAW_POWER = 0, 1
SGS_NONE, SGS_WALE = 0, 1
MAX_TERMS = 4


class Map1d(C.Structure):
    _fields_ = [("kind", C.c_int), ("p0", C.c_double), ("p1", C.c_double), ("beta", C.c_double)]


class Grid(C.Structure):
    _fields_ = [("nb", C.c_int * 3), ("nc", C.c_int * 3), ("ng", C.c_int),
                ("lo", C.c_double * 3), ("hi", C.c_double * 3), ("periodic", C.c_int * 3),
                ("map", Map1d * 3)]


class Gas(C.Structure):
    _fields_ = [("gamma", C.c_double), ("R", C.c_double)]


class Term(C.Structure):
    _fields_ = [("kind", C.c_int), ("order", C.c_int), ("sensor", C.c_int),
                ("sensor_param", C.c_double), ("eps_p", C.c_double), ("eps_T", C.c_double),
                ("split", C.c_int), ("smooth", C.c_int), ("law", C.c_int),
                ("mu", C.c_double), ("T_ref", C.c_double), ("power", C.c_double),
                ("prandtl", C.c_double), ("diss_kind", C.c_int), ("full", C.c_int),
                ("sgs", C.c_int), ("sgs_cw", C.c_double), ("sgs_delta", C.c_double), ("sgs_prt", C.c_double)]


class Flux(C.Structure):
    _fields_ = [("nterms", C.c_int), ("terms", Term * MAX_TERMS)]


class FillKernel(C.Structure):
    _fields_ = [("dom", C.c_double * 5), ("fill", C.c_double * 5), ("add", C.c_double * 5),
                ("dom_normal", C.c_double), ("fill_normal", C.c_double), ("add_normal", C.c_double)]


ORA_RK_MAX = 6


class RKTable(C.Structure):
    _fields_ = [("stages", C.c_int), ("high_storage", C.c_int),
                ("a", (C.c_long * 2) * ORA_RK_MAX * ORA_RK_MAX), ("b", (C.c_long * 2) * ORA_RK_MAX),
                ("c", (C.c_long * 2) * ORA_RK_MAX)]


# explicit Butcher tables of time-integration/explicit.h:41-116 as (num, den):
# rows a (j < i), accumulation b, stage times c, high storage
_H = (1, 2)
RK_TABLES = {
    "rk2": ([[], [_H]], [(0, 1), (1, 1)], [(0, 1), _H], False),
    "rk2hs": ([[], [_H]], [(0, 1), (1, 1)], [(0, 1), _H], True),
    "rk4": ([[], [_H], [(0, 1), _H], [(0, 1), (0, 1), (1, 1)]], [(1, 6), (1, 3), (1, 3), (1, 6)],
            [(0, 1), _H, _H, (1, 1)], False),
    "ssprk34": ([[], [_H], [_H, _H], [(1, 6), (1, 6), (1, 6)]], [(1, 6), (1, 6), (1, 6), _H],
                [(0, 1), _H, (1, 1), _H], False),
    "ssprk3": ([[], [(1, 1)], [(1, 4), (1, 4)]], [(1, 6), (1, 6), (2, 3)], [(0, 1), (1, 1), _H], False),
    "ssprk3hs": ([[], [(1, 1)], [(1, 4), (1, 4)]], [(1, 6), (1, 6), (2, 3)], [(0, 1), (1, 1), _H], True),
    "rk38r": ([[], [(1, 3)], [(-1, 3), (1, 1)], [(1, 1), (-1, 1), (1, 1)]], [(1, 8), (3, 8), (3, 8), (1, 8)],
              [(0, 1), (1, 3), (2, 3), (1, 1)], False),
}


def rk_table(name):
    a, b, c, hs = RK_TABLES[name]
    t = RKTable()
    t.stages = len(b)
    t.high_storage = int(hs)
    for i, row in enumerate(a):
        for j, (n, d) in enumerate(row):
            t.a[i][j][0], t.a[i][j][1] = n, d
    for i, (n, d) in enumerate(b):
        t.b[i][0], t.b[i][1] = n, d
    for i, (n, d) in enumerate(c):
        t.c[i][0], t.c[i][1] = n, d
    return t


class BC(C.Structure):
    _fields_ = [("wall_mask", C.c_int), ("t_wall", C.c_double)]


def _load():
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        raise FileNotFoundError(f"oracle library not built: {path} (run make -C oracle)")
    lib = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    lib.ora_array_size.argtypes = [C.POINTER(Grid)]
    lib.ora_set_threads.argtypes = [C.c_int]
    lib.ora_flux_div.argtypes = [C.POINTER(Grid), C.POINTER(Gas), C.POINTER(Flux), dp, dp, C.c_int]
    lib.ora_flux_div_blocks.argtypes = [C.POINTER(Grid), C.POINTER(Gas), C.POINTER(Flux), dp, dp, C.c_int,
                                        C.POINTER(C.c_int), C.c_int]
    lib.ora_exchange.argtypes = [C.POINTER(Grid), dp]
    lib.ora_boundary_fill_wall.argtypes = [C.POINTER(Grid), dp, C.c_int, C.c_double]
    lib.ora_rk_transform_step.argtypes = [C.POINTER(Grid), C.POINTER(Gas), C.POINTER(Flux), C.POINTER(BC),
                                          C.POINTER(RKTable), C.c_double, dp, C.POINTER(dp)]
    lib.ora_rk_identity_step.argtypes = [C.POINTER(Grid), C.POINTER(Gas), C.POINTER(Flux), C.POINTER(RKTable),
                                         C.c_double, dp, dp, C.POINTER(dp), dp]
    lib.ora_boundary_fill_extrap.argtypes = [C.POINTER(Grid), dp, C.c_int, C.c_int]
    lib.ora_boundary_fill_affine.argtypes = [C.POINTER(Grid), dp, C.c_int, C.POINTER(FillKernel)]
    lib.ora_prim_to_cons.argtypes = [C.POINTER(Gas), dp, dp, C.c_long]
    lib.ora_cons_to_prim.argtypes = [C.POINTER(Gas), dp, dp, C.c_long]
    lib.ora_ssprk3_step.argtypes = [C.POINTER(Grid), C.POINTER(Gas), C.POINTER(Flux), C.POINTER(BC),
                                    C.c_double, dp, dp, dp, dp]
    lib.ora_ssprk3opt_step.argtypes = [C.POINTER(Grid), C.POINTER(Gas), C.POINTER(Flux), C.POINTER(BC),
                                       C.c_double, dp, dp, dp]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def make_grid(nb, nc, ng, lo, hi, periodic=(1, 1, 1), maps=None):
    g = Grid()
    for d in range(3):
        g.nb[d], g.nc[d] = nb[d], nc[d]
        g.lo[d], g.hi[d] = lo[d], hi[d]
        g.periodic[d] = int(periodic[d])
        m = (maps or [None] * 3)[d]
        if m is None:
            g.map[d].kind = MAP_IDENTITY
        else:
            g.map[d].kind, g.map[d].p0, g.map[d].p1, g.map[d].beta = m
    g.ng = ng
    return g


def make_term(kind, **kw):
    t = Term()
    t.kind = kind
    for k, v in kw.items():
        setattr(t, k, v)
    return t


def make_flux(*terms):
    f = Flux()
    f.nterms = len(terms)
    for i, t in enumerate(terms):
        f.terms[i] = t
    return f


def shape_of(g):
    nblk = g.nb[0] * g.nb[1] * g.nb[2]
    return (nblk, g.nc[2] + 2 * g.ng, g.nc[1] + 2 * g.ng, g.nc[0] + 2 * g.ng, 5)


def _ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def set_threads(n):
    """Threads of the oracle's OpenMP loop over blocks; returns the count in use."""
    return lib().ora_set_threads(int(n))


def flux_div(g, gas, flux, prim, rhs=None, increment=False):
    if rhs is None:
        rhs = np.zeros(shape_of(g))
    rc = lib().ora_flux_div(C.byref(g), C.byref(gas), C.byref(flux), _ptr(prim), _ptr(rhs), int(increment))
    if rc != 0:
        raise RuntimeError("ora_flux_div failed")
    return rhs


def flux_div_blocks(g, gas, flux, prim, blocks, rhs=None, increment=False):
    """flux_div on the interior cells of the listed blocks only."""
    if rhs is None:
        rhs = np.zeros(shape_of(g))
    bl = np.ascontiguousarray(blocks, dtype=np.int32)
    rc = lib().ora_flux_div_blocks(C.byref(g), C.byref(gas), C.byref(flux), _ptr(prim), _ptr(rhs), int(increment),
                                   bl.ctypes.data_as(C.POINTER(C.c_int)), len(bl))
    if rc != 0:
        raise RuntimeError("ora_flux_div_blocks failed")
    return rhs


def exchange(g, a):
    lib().ora_exchange(C.byref(g), _ptr(a))
    return a


def boundary_fill_wall(g, a, face, t_wall):
    lib().ora_boundary_fill_wall(C.byref(g), _ptr(a), face, t_wall)
    return a


_dp = C.POINTER(C.c_double)


def rk_transform_step(g, gas, flux, bc, table, dt, q, steps=1):
    """Explicit table with the prim<->cons transform (advance.h:233-280)."""
    res = [np.zeros_like(q) for _ in range(table.stages)]
    ptrs = (_dp * table.stages)(*[_ptr(r) for r in res])
    for _ in range(steps):
        if lib().ora_rk_transform_step(C.byref(g), C.byref(gas), C.byref(flux), C.byref(bc or BC()), C.byref(table),
                                       dt, _ptr(q), ptrs) != 0:
            raise RuntimeError("ora_rk_transform_step failed")
    return q


def rk_identity_step(g, gas, flux, table, dt, q, steps=1):
    """Explicit table with the identity transform on conserved storage (advance.h:107-230)."""
    res = [np.zeros_like(q) for _ in range(table.stages)]
    ptrs = (_dp * table.stages)(*[_ptr(r) for r in res])
    q2, prim = np.zeros_like(q), np.zeros_like(q)
    for _ in range(steps):
        if lib().ora_rk_identity_step(C.byref(g), C.byref(gas), C.byref(flux), C.byref(table), dt, _ptr(q),
                                      _ptr(q2), ptrs, _ptr(prim)) != 0:
            raise RuntimeError("ora_rk_identity_step failed")
    return q


def boundary_fill_extrap(g, a, face_mask, order):
    if lib().ora_boundary_fill_extrap(C.byref(g), _ptr(a), face_mask, order) != 0:
        raise RuntimeError("ora_boundary_fill_extrap failed")
    return a


def fill_kernel(dom, fill=(0.0,) * 5, add=(0.0,) * 5, dom_normal=None, fill_normal=None, add_normal=None):
    """Affine boundary kernel: ghost[v] = dom[v] image + fill[v] ghost + add[v];
    the face-normal velocity takes dom_normal / fill_normal / add_normal,
    which default to the u entries (index 2)."""
    k = FillKernel()
    for v in range(5):
        k.dom[v], k.fill[v], k.add[v] = dom[v], fill[v], add[v]
    k.dom_normal = dom[2] if dom_normal is None else dom_normal
    k.fill_normal = fill[2] if fill_normal is None else fill_normal
    k.add_normal = add[2] if add_normal is None else add_normal
    return k


def boundary_fill_affine(g, a, face_mask, kern):
    if lib().ora_boundary_fill_affine(C.byref(g), _ptr(a), face_mask, C.byref(kern)) != 0:
        raise RuntimeError("ora_boundary_fill_affine failed")
    return a


def prim_to_cons(gas, prim):
    out = np.empty_like(prim)
    lib().ora_prim_to_cons(C.byref(gas), _ptr(prim), _ptr(out), prim.size // 5)
    return out


def cons_to_prim(gas, cons):
    out = np.empty_like(cons)
    lib().ora_cons_to_prim(C.byref(gas), _ptr(cons), _ptr(out), cons.size // 5)
    return out


def ssprk3_step(g, gas, flux, bc, dt, q, nsteps=1):
    r = [np.zeros(shape_of(g)) for _ in range(3)]
    for _ in range(nsteps):
        lib().ora_ssprk3_step(C.byref(g), C.byref(gas), C.byref(flux), C.byref(bc) if bc else None,
                              dt, _ptr(q), _ptr(r[0]), _ptr(r[1]), _ptr(r[2]))
    return q


def ssprk3opt_step(g, gas, flux, bc, dt, q, nsteps=1):
    r = [np.zeros(shape_of(g)) for _ in range(2)]
    for _ in range(nsteps):
        lib().ora_ssprk3opt_step(C.byref(g), C.byref(gas), C.byref(flux), C.byref(bc) if bc else None,
                                 dt, _ptr(q), _ptr(r[0]), _ptr(r[1]))
    return q
