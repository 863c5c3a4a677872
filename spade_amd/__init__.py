"""spade_amd: MI355X-native right-hand-side hot path of SPADE.

Python mirror of the reference's C++ interface for this path, calling the C
ABI of libspade_amd.so (include/spade_amd.h):

  reference (SPADE src/)                         here
  ---------------------------------------------  ----------------------------------
  cartesian_blocks_t + cartesian_grid_t          CartesianGrid
  (grid/cartesian_blocks.h, cartesian_grid.h)
  grid_array (grid/grid_array.h)                 GridArray
  ideal_gas_t (navier-stokes/gas.h)              IdealGas
  cent_keep<order>, totani_lr, pressure_diss_lr, cent_keep, totani_lr, pressure_diss_lr,
  weno_t, fweno_t, weno_fds_t, first_order_t,    weno, fweno, weno_fds, first_order,
  visc_lr                                        visc_lr
  (navier-stokes/convective.h, viscous.h)
  constant_viscosity_t, power_law_t              constant_viscosity, power_law
  const_sensor_t, ducros_t                       const_sensor, ducros
  omni::compose                                  compose
  pde_algs::flux_div(+ overwrite/increment)      flux_div(prims, rhs, flux, increment=True)
  create_exchange(...).exchange(arr)             create_exchange(arr).exchange(arr)
  algs::boundary_fill (wall mirror)              boundary_fill_wall
  integrator_t(...).advance() with ssprk3_t /    Integrator(q, flux, bc, scheme, dt).advance()
  ssprk3_opt and state_transform_t
  io::output_vtk(dir, name, arr) (io/io_vtk.h)   output_vtk(dir, name, arr[, names])
  io::binary_write / binary_read (io_native.h)   binary_write(file, arr) / binary_read(file, arr)

Every array lives on the GPU; there is no CPU fallback.  If the library is
missing, importing works but the first call raises FileNotFoundError.
"""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import SpadeError, check, load

__all__ = [
    "Context", "CartesianGrid", "GridArray", "IdealGas", "cent_keep", "totani_lr", "pressure_diss_lr",
    "weno", "fweno", "weno_fds", "first_order", "visc_lr", "constant_viscosity", "power_law", "const_sensor",
    "ducros", "compose", "flux_div", "create_exchange", "boundary_fill_wall", "boundary_fill", "extrapolate",
    "FillKernel", "isothermal_wall", "xmin", "xmax", "ymin", "ymax", "zmin", "zmax", "Integrator",
    "ssprk3", "ssprk3_opt", "max_wavespeed_noalloc", "RKIntegrator", "rk_table", "rk_stage", "rk2", "rk2hs", "rk4", "ssprk34",
    "ssprk3_table", "ssprk3hs", "rk38r", "prim_cons", "identity_transform", "max_wavespeed", "tanh_map", "SpadeError", "diagnostics", "CaseSpec", "Batch",
    "grid_desc", "output_vtk", "output_vtk_host", "binary_write", "binary_read", "PRIM_NAMES", "CONS_NAMES",
    "FLUX_NAMES",
]

# ---------------------------------------------------------------- context ---


class Context:
    """One GPU and one HIP stream; every operation is ordered on that stream."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        check(load().spd_ctx_create(device, C.byref(self._h)))
        self.device = device

    def sync(self):
        check(load().spd_ctx_sync(self._h))

    @property
    def stream(self):
        return load().spd_ctx_stream(self._h)

    def timer_start(self):
        check(load().spd_timer_start(self._h))

    def timer_stop(self):
        ms = C.c_float()
        check(load().spd_timer_stop(self._h, C.byref(ms)))
        return ms.value

    def close(self):
        if self._h:
            load().spd_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


# ------------------------------------------------------------------- grid ---


def tanh_map(p0, p1, beta):
    """1-D stretching x = c + h tanh(beta s)/tanh(beta), s = (xi - c)/h on [p0, p1]."""
    return (_abi.MAP_TANH, float(p0), float(p1), float(beta))


def grid_desc(cells_in_block, num_blocks, bounds, num_exch=2, periodic=(True, True, True),
              maps=(None, None, None)):
    """The C ABI's grid description (spd_grid_desc) of a CartesianGrid with these
    arguments; needs no device."""
    d = _abi.GridDesc()
    for a in range(3):
        d.num_blocks[a] = int(num_blocks[a])
        d.cells_in_block[a] = int(cells_in_block[a])
        d.lo[a], d.hi[a] = float(bounds[a][0]), float(bounds[a][1])
        d.periodic[a] = int(bool(periodic[a]))
        m = maps[a]
        if m is None:
            d.map[a].kind = _abi.MAP_IDENTITY
        else:
            d.map[a].kind, d.map[a].p0, d.map[a].p1, d.map[a].beta = m
    d.num_exch = int(num_exch)
    return d


class CartesianGrid:
    """Uniform multi-block grid (cartesian_blocks_t + cartesian_grid_t).

    cells_in_block, num_blocks: per-axis counts; bounds: ((x0,x1),(y0,y1),(z0,z1))
    in computational coordinates; maps: per-axis None (identity) or tanh_map().
    """

    def __init__(self, cells_in_block, num_blocks, bounds, num_exch=2, periodic=(True, True, True),
                 maps=(None, None, None), ctx=None):
        self.ctx = ctx or default_context()
        d = grid_desc(cells_in_block, num_blocks, bounds, num_exch, periodic, maps)
        self.desc = d
        self._h = C.c_void_p()
        check(load().spd_grid_create(self.ctx._h, C.byref(d), C.byref(self._h)))
        self.cells_in_block = tuple(int(c) for c in cells_in_block)
        self.num_blocks = tuple(int(b) for b in num_blocks)
        self.num_exch = int(num_exch)
        self.periodic = tuple(bool(p) for p in periodic)

    @property
    def nblocks(self):
        return self.num_blocks[0] * self.num_blocks[1] * self.num_blocks[2]

    @property
    def num_cells(self):
        c = self.cells_in_block
        return c[0] * c[1] * c[2] * self.nblocks

    @property
    def device_size(self):
        """Doubles one array of this grid occupies on the device (blocks padded to 256 bytes)."""
        return int(load().spd_array_device_size(self._h))

    @property
    def host_shape(self):
        """Shape of the reference-layout host image: (blocks, Nz, Ny, Nx, 5)."""
        ng = self.num_exch
        c = self.cells_in_block
        return (self.nblocks, c[2] + 2 * ng, c[1] + 2 * ng, c[0] + 2 * ng, 5)

    def close(self):
        if self._h:
            load().spd_grid_destroy(self._h)
            self._h = C.c_void_p()


class GridArray:
    """5-variable FP64 cell-centred array with ghosts (grid_array)."""

    def __init__(self, grid, fill=None):
        self.grid = grid
        self._h = C.c_void_p()
        check(load().spd_array_create(grid._h, C.byref(self._h)))
        if fill is not None:
            self.fill(fill)

    def upload(self, host):
        a = np.ascontiguousarray(host, dtype=np.float64)
        if a.shape != self.grid.host_shape:
            raise ValueError(f"host array shape {a.shape} != {self.grid.host_shape}")
        check(load().spd_array_upload(self._h, a.ctypes.data_as(C.POINTER(C.c_double))))
        return self

    def download(self):
        out = np.empty(self.grid.host_shape, dtype=np.float64)
        check(load().spd_array_download(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def fill(self, value):
        check(load().spd_array_fill(self._h, float(value)))
        return self

    def copy_from(self, other):
        check(load().spd_array_copy(self._h, other._h))
        return self

    # grid_array element-wise operations (grid/grid_array.h:286-396)
    def __imul__(self, s):
        check(load().spd_array_scale(self._h, float(s)))
        return self

    def __iadd__(self, other):
        check(load().spd_array_add(self._h, other._h))
        return self

    def __isub__(self, other):
        check(load().spd_array_sub(self._h, other._h))
        return self

    def convert(self, direction, gas, include_ghosts=True):
        """convert_state in place: direction "cons" (prim -> cons) or "prim"."""
        d = {"cons": _abi.TO_CONS, "prim": _abi.TO_PRIM}[direction]
        check(load().spd_array_convert(self._h, d, gas.gamma, gas.R, int(bool(include_ghosts))))
        return self

    @classmethod
    def _view(cls, grid, handle):
        """A non-owning wrapper of an array handle the library passes to a callback."""
        v = cls.__new__(cls)
        v.grid = grid
        v._h = C.c_void_p(handle)
        v._view_only = True
        return v

    def close(self):
        if self._h and not getattr(self, "_view_only", False):
            load().spd_array_destroy(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- physics ---


class IdealGas:
    def __init__(self, gamma, R):
        self.gamma, self.R = float(gamma), float(R)


class _Term:
    def __init__(self, gas, **fields):
        self.gas = gas
        self.fields = fields

    def terms(self):
        return [self]


def cent_keep(order, gas):
    return _Term(gas, kind=_abi.TERM_KEEP, order=int(order))


def totani_lr(gas):
    return _Term(gas, kind=_abi.TERM_TOTANI)


class _Sensor:
    def __init__(self, kind, param):
        self.kind, self.param = kind, float(param)


def const_sensor(alpha):
    return _Sensor(_abi.SENSOR_CONST, alpha)


def ducros(epsilon):
    return _Sensor(_abi.SENSOR_DUCROS, epsilon)


def pressure_diss_lr(gas, sensor, eps_p, eps_T=None):
    return _Term(gas, kind=_abi.TERM_PDISS, sensor=sensor.kind, sensor_param=sensor.param,
                 eps_p=float(eps_p), eps_T=float(eps_p if eps_T is None else eps_T))


def weno(gas, split="rusanov", smooth=True):
    return _Term(gas, kind=_abi.TERM_WENO, split=_abi.SPLIT_PHYS if split == "phys" else _abi.SPLIT_RUSANOV,
                 smooth=int(bool(smooth)))


def fweno(gas, smooth=True):
    return _Term(gas, kind=_abi.TERM_FWENO, smooth=int(bool(smooth)))


def first_order(gas):
    return _Term(gas, kind=_abi.TERM_FIRST)


def weno_fds(gas, smooth=True):
    """weno_fds_t<rusanov_fds_t> (navier-stokes/convective.h:499-555,
    flux_funcs.h:54-95).  With smooth=True (enable_smooth) the reference's
    face states are both the central value, so its flux is identically zero."""
    return _Term(gas, kind=_abi.TERM_WENO_FDS, smooth=int(bool(smooth)))


# hybrid_scheme_t flux tags (navier-stokes/hybrid_scheme.h:7-12)
full_flux, diss_flux = 1, 0


def hybrid_scheme(central, dissipative, sensor, flux_tag=full_flux):
    """hybrid_scheme_t(scheme0, scheme1, blender, tag) (navier-stokes/hybrid_scheme.h:14-48):
    cent_keep<2|4> blended with weno(...) or fweno(...) by the face sensor b
    (const_sensor or ducros); full_flux = (1 - b) F0 + b F1, diss_flux = F0 + b F1.
    The descriptor is well formed, but flux_div refuses it (SpadeError, code
    SPD_ERR_UNSUPPORTED): the device library has no code for this scheme."""
    c, d = central.terms(), dissipative.terms()
    if len(c) != 1 or c[0].fields.get("kind") != _abi.TERM_KEEP:
        raise ValueError("hybrid_scheme: the central scheme must be cent_keep(order, gas)")
    if len(d) != 1 or d[0].fields.get("kind") not in (_abi.TERM_WENO, _abi.TERM_FWENO):
        raise ValueError("hybrid_scheme: the dissipative scheme must be weno(...) or fweno(...)")
    if c[0].gas.gamma != d[0].gas.gamma or c[0].gas.R != d[0].gas.R:
        raise ValueError("hybrid_scheme: both schemes must use the same gas")
    df = d[0].fields
    return _Term(c[0].gas, kind=_abi.TERM_HYBRID, order=c[0].fields["order"], diss_kind=df["kind"],
                 split=df.get("split", _abi.SPLIT_RUSANOV), smooth=df.get("smooth", 1), sensor=sensor.kind,
                 sensor_param=sensor.param, full=int(flux_tag))


class _Law:
    def __init__(self, **fields):
        self.fields = fields


def constant_viscosity(mu, prandtl):
    return _Law(law=_abi.LAW_CONST, mu=float(mu), prandtl=float(prandtl))


def power_law(mu_ref, T_ref, power, prandtl):
    return _Law(law=_abi.LAW_POWER, mu=float(mu_ref), T_ref=float(T_ref), power=float(power),
                prandtl=float(prandtl))


def wale(gas, cw, delta, prt):
    """wale_t(gas, cw, delta, prt) (subgrid_scale.h:24-86): the WALE subgrid
    viscosity with coefficient cw, filter width delta and turbulent Prandtl
    number prt; combine with a laminar law through sgs_visc."""
    return dict(sgs=_abi.SGS_WALE, sgs_cw=float(cw), sgs_delta=float(delta), sgs_prt=float(prt))


def sgs_visc(law, model):
    """sgs_visc_t(laminar law, subgrid model) (viscous_laws.h:174-216)."""
    return _Law(**law.fields, **model)


def visc_lr(law, gas):
    return _Term(gas, kind=_abi.TERM_VISC, **law.fields)


class _Composite:
    def __init__(self, parts):
        self.parts = parts

    def terms(self):
        out = []
        for p in self.parts:
            out.extend(p.terms())
        return out


def compose(*fluxes):
    """F0 + (F1 + ...) evaluated per face (omni::compose)."""
    return _Composite(list(fluxes))


def flux_descriptor(flux):
    terms = flux.terms()
    if not 1 <= len(terms) <= _abi.MAX_TERMS:
        raise ValueError("a flux needs 1..4 terms")
    gas = terms[0].gas
    d = _abi.FluxDesc()
    d.nterms = len(terms)
    for i, t in enumerate(terms):
        if t.gas.gamma != gas.gamma or t.gas.R != gas.R:
            raise ValueError("all composed fluxes must use the same gas")
        for k, v in t.fields.items():
            setattr(d.terms[i], k, v)
    d.gamma, d.R = gas.gamma, gas.R
    return d


# ------------------------------------------------------------- algorithms ---


def flux_div(prims, rhs, flux, increment=True):
    """rhs (+)= -div F over interior cells (pde_algs::flux_div; default increment)."""
    d = flux_descriptor(flux)
    check(load().spd_flux_div(prims._h, rhs._h, C.byref(d), _abi.INCREMENT if increment else _abi.OVERWRITE))


def flux_div_kernel(prims, flux):
    """'fast' (LDS face-once kernel), 'specialized' (per-cell kernel compiled for
    a composition of cent_keep<2>, pressure_diss_lr and visc_lr) or 'generic'
    (per-cell kernel for every functor) for this array and flux."""
    d = flux_descriptor(flux)
    kind = C.c_int()
    check(load().spd_flux_div_kernel(prims._h, C.byref(d), C.byref(kind)))
    return {1: "fast", 2: "specialized"}.get(kind.value, "generic")


class ExchangeHandle:
    def exchange(self, arr):
        check(load().spd_exchange(arr._h))


def create_exchange(arr, periodic=None):
    """Ghost exchange handle; periodicity is a property of the grid here."""
    if periodic is not None and tuple(bool(p) for p in periodic) != arr.grid.periodic:
        raise ValueError("periodicity is fixed when the grid is created")
    return ExchangeHandle()


def boundary_fill_wall(arr, face, t_wall):
    """Isothermal no-slip mirror ghosts on face = 2*dir + (0 min | 1 max)."""
    check(load().spd_boundary_fill_wall(arr._h, int(face), float(t_wall)))


# boundary identifiers (grid/boundary_fill.h:9-15), combined with |
xmin, xmax, ymin, ymax, zmin, zmax = 1, 2, 4, 8, 16, 32


class extrapolate:
    """boundary::extrapolate<order> (grid/boundary_fill.h:17-26)."""

    def __init__(self, order):
        self.order = int(order)


class FillKernel:
    """A boundary_fill kernel kern(q_image, q_ghost, x, idir) given as data:
    per primitive v (p, T, u, v, w), ghost = dom[v] q_image + fill[v] q_ghost
    + add[v]; the velocity component normal to the face takes dom_normal /
    fill_normal / add_normal (default: the u entries).  Zero coefficients
    leave their term out (grid/boundary_fill.h:99-127)."""

    def __init__(self, dom, fill=(0.0,) * 5, add=(0.0,) * 5, dom_normal=None, fill_normal=None, add_normal=None):
        k = _abi.FillKernel()
        for v in range(5):
            k.dom[v], k.fill[v], k.add[v] = float(dom[v]), float(fill[v]), float(add[v])
        k.dom_normal = float(dom[2] if dom_normal is None else dom_normal)
        k.fill_normal = float(fill[2] if fill_normal is None else fill_normal)
        k.add_normal = float(add[2] if add_normal is None else add_normal)
        self._k = k


def isothermal_wall(t_wall):
    """No-slip isothermal mirror: (p, 2 T_wall - T, -u, -v, -w) of the image cell."""
    return FillKernel(dom=(1.0, -1.0, -1.0, -1.0, -1.0), add=(0.0, 2.0 * t_wall, 0.0, 0.0, 0.0))


def boundary_fill(arr, faces, kernel):
    """algs::boundary_fill(arr, boundaries, kern) (grid/boundary_fill.h:32-133):
    faces = xmin | ymax | ..., filled in face order; kernel = extrapolate(order)
    or a FillKernel."""
    if isinstance(kernel, extrapolate):
        check(load().spd_boundary_fill_extrap(arr._h, int(faces), kernel.order))
    elif isinstance(kernel, FillKernel):
        check(load().spd_boundary_fill(arr._h, int(faces), C.byref(kernel._k)))
    else:
        raise TypeError("boundary_fill kernel must be extrapolate(order) or FillKernel(...)")


ssprk3 = _abi.SCHEME_SSPRK3
ssprk3_opt = _abi.SCHEME_SSPRK3_OPT


class Integrator:
    """integrator_t with rhs = flux_div (overwrite), bc = exchange + walls and the
    prim<->cons state transform; the solution array q is advanced in place."""

    def __init__(self, q, flux, dt, scheme=ssprk3, walls=(), t_wall=0.0):
        self.q = q
        self.dt = float(dt)
        bc = _abi.BCDesc()
        bc.wall_mask = 0
        for face in walls:
            bc.wall_mask |= 1 << int(face)
        bc.t_wall = float(t_wall)
        self._d = flux_descriptor(flux)
        self._h = C.c_void_p()
        check(load().spd_integrator_create(q._h, C.byref(self._d), C.byref(bc), int(scheme), self.dt,
                                           C.byref(self._h)))

    def advance(self, nsteps=1):
        check(load().spd_integrator_advance(self._h, int(nsteps)))

    @property
    def fused(self):
        """True when each SSP-RK3 stage runs as one fused flux_div + stage launch
        (same bits as the separate launches)."""
        f = C.c_int()
        check(load().spd_integrator_fused(self._h, C.byref(f)))
        return bool(f.value)

    def solution(self):
        return self.q

    def close(self):
        if self._h:
            load().spd_integrator_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# explicit Butcher tables (time-integration/explicit.h:41-116)
rk2, rk2hs, rk4, ssprk34, ssprk3_table, ssprk3hs, rk38r = (_abi.RK2, _abi.RK2HS, _abi.RK4, _abi.SSPRK34,
                                                         _abi.SSPRK3_TABLE, _abi.SSPRK3HS, _abi.RK38R)
# state transforms: state_transform_t(cons_t, gas) and identity_transform
prim_cons, identity_transform = _abi.TRANS_PRIM_CONS, _abi.TRANS_IDENTITY


def rk_table(which=None, a=None, b=None, c=None, high_storage=False):
    """A built-in table (rk4, ...) or one given as lists of (num, den) ratios:
    a[i][j] for j < i, b, c -- the rk_t<table, accum, dt> of explicit.h:27-38."""
    t = _abi.RKTable()
    if which is not None:
        check(load().spd_rk_table_builtin(int(which), C.byref(t)))
        return t
    t.stages = len(b)
    t.high_storage = int(bool(high_storage))
    for i, row in enumerate(a):
        for j, (n, d) in enumerate(row):
            t.a[i][j].num, t.a[i][j].den = n, d
    for i, (n, d) in enumerate(b):
        t.b[i].num, t.b[i].den = n, d
    for i, (n, d) in enumerate(c):
        t.c[i].num, t.c[i].den = n, d
    return t


class RKIntegrator:
    """integrator_t(axis, scheme, data, rhs, bc, trans) with any explicit table
    (time-integration/integrator.h:18-59): trans = prim_cons (the solution holds
    primitives; transform-aware advance, advance.h:233-280) or identity_transform
    (generic advance on the stored variables, advance.h:107-230).  rhs(resid,
    sol, t) and bc(sol, t) are Python callables on GridArray views; rhs = None is
    flux_div(sol, resid, flux) (prim_cons only), bc = None the exchange plus the
    isothermal walls."""

    def __init__(self, q, table, dt, t0=0.0, trans=prim_cons, gas=None, flux=None, rhs=None, bc=None,
                 walls=(), t_wall=0.0):
        self.q = q
        d = _abi.RKDesc()
        d.table = rk_table(table) if isinstance(table, int) else table
        d.trans = int(trans)
        if gas is None and flux is not None:
            gas = flux.terms()[0].gas
        if gas is not None:
            d.gamma, d.R = gas.gamma, gas.R
        d.t0, d.dt = float(t0), float(dt)
        self._errors = []
        grid = q.grid

        def guard(fn):
            def call(*args):
                try:
                    fn(*args)
                    return 0
                except Exception as e:            # reported by advance()
                    self._errors.append(e)
                    return 1
            return call
        if rhs is not None:
            self._rhs = _abi.RHS_FN(guard(lambda r, s, t, u: rhs(GridArray._view(grid, r), GridArray._view(grid, s), t)))
            d.rhs = self._rhs
        if bc is not None:
            self._bc = _abi.BC_FN(guard(lambda s, t, u: bc(GridArray._view(grid, s), t)))
            d.bc = self._bc
        if flux is not None:
            self._fd = flux_descriptor(flux)
            d.flux = C.pointer(self._fd)
        if walls:
            self._walls = _abi.BCDesc()
            for face in walls:
                self._walls.wall_mask |= 1 << int(face)
            self._walls.t_wall = float(t_wall)
            d.walls = C.pointer(self._walls)
        self._d = d
        self._h = C.c_void_p()
        check(load().spd_rk_integrator_create(q._h, C.byref(d), C.byref(self._h)))

    def advance(self, nsteps=1):
        rc = load().spd_integrator_advance(self._h, int(nsteps))
        if self._errors:
            e, self._errors = self._errors[0], []
            raise e
        check(rc)

    def time(self):
        t = C.c_double()
        check(load().spd_integrator_time(self._h, C.byref(t)))
        return t.value

    def solution(self):
        return self.q

    def close(self):
        if self._h:
            load().spd_integrator_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rk_stage(prims, resids, coefs, gas, push_ghosts=False):
    """transform_advance_to (advance.h:56-102): cons = p2c(q); cons += coef_j r_j;
    q = c2p(cons) on the interior cells; coefs are the caller's (a_ij - a_{i-1,j}) dt."""
    n = len(resids)
    hs = (C.c_void_p * n)(*[r._h.value for r in resids])
    cs = (C.c_double * n)(*[float(c) for c in coefs])
    check(load().spd_rk_stage(prims._h, hs, cs, n, gas.gamma, gas.R, int(bool(push_ghosts))))


def diagnostics(prims, gas):
    """(volume-averaged kinetic energy, enstrophy) of a primitive array with current ghosts."""
    ke, en = C.c_double(), C.c_double()
    check(load().spd_diagnostics(prims._h, gas.R, C.byref(ke), C.byref(en)))
    return ke.value, en.value


class CaseSpec:
    """One case of the batched engine (TGV or isentropic vortex, periodic)."""

    def __init__(self, kind="tgv", cells=(64, 64, 64), mach=0.1, reynolds=1600.0, gamma=1.4, prandtl=0.72,
                 law="power", power=0.76, cfl=0.5, steps=10):
        self.kind, self.cells = kind, tuple(int(c) for c in cells)
        self.mach, self.reynolds, self.gamma, self.prandtl = float(mach), float(reynolds), float(gamma), float(prandtl)
        self.law, self.power, self.cfl, self.steps = law, float(power), float(cfl), int(steps)

    @property
    def blocks(self):
        return (self.cells[0] // 16) * (self.cells[1] // 16) * (self.cells[2] // 16)

    @property
    def num_cells(self):
        return self.cells[0] * self.cells[1] * self.cells[2]

    def desc(self):
        d = _abi.CaseDesc()
        d.kind = _abi.CASE_TGV if self.kind == "tgv" else _abi.CASE_VORTEX
        for a in range(3):
            d.cells[a] = self.cells[a]
        d.mach, d.reynolds, d.gamma, d.prandtl = self.mach, self.reynolds, self.gamma, self.prandtl
        d.law = _abi.LAW_POWER if self.law == "power" else _abi.LAW_CONST
        d.power, d.cfl = self.power, self.cfl
        return d


class Batch:
    """Batched FP32-flux / FP64-state engine (device pool sized by an HBM budget)."""

    BYTES_PER_BLOCK = 3 * 5 * 4096 * 8

    def __init__(self, hbm_budget_bytes, law="power", ctx=None):
        self.ctx = ctx or default_context()
        self._h = C.c_void_p()
        lawi = _abi.LAW_POWER if law == "power" else _abi.LAW_CONST
        check(load().spd_batch_create(self.ctx._h, int(hbm_budget_bytes), lawi, C.byref(self._h)))

    def capacity(self, spec):
        need, free = C.c_longlong(), C.c_longlong()
        d = spec.desc()
        check(load().spd_batch_capacity(self._h, C.byref(d), C.byref(need), C.byref(free)))
        return need.value, free.value

    def fits(self, spec):
        need, free = self.capacity(spec)
        return need <= free

    def add(self, spec):
        d = spec.desc()
        cid = C.c_int()
        check(load().spd_batch_add_case(self._h, C.byref(d), C.byref(cid)))
        return cid.value

    def remove(self, cid):
        check(load().spd_batch_remove_case(self._h, int(cid)))

    def step(self, nsteps=1):
        check(load().spd_batch_step(self._h, int(nsteps)))

    def info(self, cid):
        cells, steps, dt = C.c_longlong(), C.c_int(), C.c_double()
        check(load().spd_batch_case_info(self._h, int(cid), C.byref(cells), C.byref(steps), C.byref(dt)))
        return {"cells": cells.value, "steps_done": steps.value, "dt": dt.value}

    def status(self, cid):
        """-1 while the case's kinetic energy stays finite, else the first step
        whose starting state had a non-finite one."""
        s = C.c_int()
        check(load().spd_batch_case_status(self._h, int(cid), C.byref(s)))
        return s.value

    def download(self, cid):
        n = self.info(cid)["cells"]
        out = np.empty(5 * n, dtype=np.float64)
        check(load().spd_batch_case_download(self._h, int(cid), out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.reshape(n // 4096, 5, 16, 16, 16)

    def upload(self, cid, cons):
        a = np.ascontiguousarray(cons, dtype=np.float64)
        check(load().spd_batch_case_upload(self._h, int(cid), a.ctypes.data_as(C.POINTER(C.c_double))))

    def diag(self, cid):
        steps = self.info(cid)["steps_done"]
        ke = np.empty(max(steps, 1)); en = np.empty(max(steps, 1))
        n = C.c_int()
        check(load().spd_batch_case_diag(self._h, int(cid), ke.ctypes.data_as(C.POINTER(C.c_double)),
                                         en.ctypes.data_as(C.POINTER(C.c_double)), steps, C.byref(n)))
        return ke[:n.value].copy(), en[:n.value].copy()

    def tune(self):
        g, inter = C.c_int(), C.c_int()
        check(load().spd_batch_tune(self._h, C.byref(g), C.byref(inter)))
        return g.value, bool(inter.value)

    def use_graph(self, enable=True):
        check(load().spd_batch_use_graph(self._h, int(bool(enable))))

    def set_shape(self, grid, interleaved=False):
        check(load().spd_batch_set_shape(self._h, int(grid), int(bool(interleaved))))

    def set_kernel_shape(self, rows=16, prefetch=1, occupancy=3):
        """Stage kernel shape: rows of a block's cross-section per workgroup
        (16, 8), planes of state in flight (1, 2) and waves per SIMD its
        registers are sized for (3, 2)."""
        check(load().spd_batch_set_kernel_shape(self._h, int(rows), int(prefetch), int(occupancy)))

    def tune_log(self):
        """Candidates timed by the last tune(): list of dicts, in timing order."""
        buf = (_abi.TuneEntry * 256)()
        n = C.c_int()
        check(load().spd_batch_tune_log(self._h, C.cast(buf, C.c_void_p), 256, C.byref(n)))
        return [{"grid": e.grid, "interleaved": bool(e.interleaved), "rows": e.rows, "prefetch": e.prefetch,
                 "occupancy": e.occupancy, "ms_per_step": e.ms_per_step} for e in buf[:n.value]]

    def shape(self):
        """The launch shape in use: grid, interleaved, rows, prefetch, occupancy."""
        v = [C.c_int() for _ in range(5)]
        check(load().spd_batch_get_shape(self._h, *[C.byref(x) for x in v]))
        return {"grid": v[0].value, "interleaved": bool(v[1].value), "rows": v[2].value,
                "prefetch": v[3].value, "occupancy": v[4].value}

    def close(self):
        if self._h:
            load().spd_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def max_wavespeed(prims, gas):
    out = C.c_double()
    check(load().spd_max_wavespeed(prims._h, gas.gamma, gas.R, C.byref(out)))
    return out.value


def max_wavespeed_noalloc(prims, gas, d_out):
    """The CFL wave speed written to d_out (address of one double of device
    memory on the array's device), stream-ordered, without allocating."""
    check(load().spd_max_wavespeed_noalloc(prims._h, gas.gamma, gas.R, C.c_void_p(int(d_out))))


# --------------------------------------------------------------------- io ---

# variable names of the reference's state types (navier-stokes/fluid_state.h:28-75)
PRIM_NAMES = ("P", "T", "U", "V", "W")
CONS_NAMES = ("rho", "rhoH", "rhoU", "rhoV", "rhoW")
FLUX_NAMES = ("continuity", "energy", "x_momentum", "y_momentum", "z_momentum")


def _names(names):
    if names is None:
        return None
    names = [str(n) for n in names]
    if len(names) != 5:
        raise ValueError("output_vtk needs 5 variable names")
    arr = (C.c_char_p * 5)(*[n.encode() for n in names])
    return C.cast(arr, C.POINTER(C.c_char_p)), arr


def output_vtk(out_dir, basename, arr, names=None):
    """io::output_vtk(out_dir, basename, array) (io/io_vtk.h:276-288):
    <out_dir>/<basename>.pvts + <out_dir>/data_<basename>/b<lb>.vts.
    names: the 5 variable names (default PRIM_NAMES, the prim_t names)."""
    n = _names(names)
    check(load().spd_io_output_vtk(arr._h, str(out_dir).encode(), str(basename).encode(), n[0] if n else None))


def output_vtk_host(out_dir, basename, desc, host, names=None):
    """output_vtk of an array image in host memory (reference layout, shape
    (blocks, Nz, Ny, Nx, 5)) on the grid `desc` (grid_desc(...)); needs no device."""
    a = np.ascontiguousarray(host, dtype=np.float64)
    ng = desc.num_exch
    nb = desc.num_blocks[0] * desc.num_blocks[1] * desc.num_blocks[2]
    shape = (nb,) + tuple(desc.cells_in_block[a_] + 2 * ng for a_ in (2, 1, 0)) + (5,)
    if a.shape != shape:
        raise ValueError(f"host array shape {a.shape} != {shape}")
    n = _names(names)
    check(load().spd_io_output_vtk_host(C.byref(desc), a.ctypes.data_as(C.POINTER(C.c_double)),
                                        str(out_dir).encode(), str(basename).encode(), n[0] if n else None))


def binary_write(filename, arr):
    """io::binary_write(filename, array) (io/io_native.h:39-45): the array's
    host image (reference layout, ghosts included), block after block."""
    check(load().spd_io_binary_write(arr._h, str(filename).encode()))


def binary_read(filename, arr):
    """io::binary_read(filename, array) (io/io_native.h:47-55)."""
    check(load().spd_io_binary_read(arr._h, str(filename).encode()))
