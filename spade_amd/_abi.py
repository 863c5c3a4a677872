"""ctypes declarations of the libspade_amd C ABI (include/spade_amd.h).

The library is loaded from this package folder only; a missing library is an
error, never a silent fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspade_amd.so")

SPD_OK, SPD_ERR_ARG, SPD_ERR_HIP, SPD_ERR_UNSUPPORTED, SPD_ERR_MEMORY, SPD_ERR_IO = range(6)
MAP_IDENTITY, MAP_TANH = 0, 1
TERM_KEEP, TERM_TOTANI, TERM_PDISS, TERM_WENO, TERM_FWENO, TERM_FIRST, TERM_VISC, TERM_HYBRID, TERM_WENO_FDS = \
    range(1, 10)
SENSOR_CONST, SENSOR_DUCROS = 0, 1
SPLIT_RUSANOV, SPLIT_PHYS = 0, 1
LAW_CONST, LAW_POWER = 0, 1
SGS_NONE, SGS_WALE = 0, 1
MAX_TERMS = 4
OVERWRITE, INCREMENT = 0, 1
SCHEME_SSPRK3, SCHEME_SSPRK3_OPT = 0, 1


class Map1d(C.Structure):
    _fields_ = [("kind", C.c_int), ("p0", C.c_double), ("p1", C.c_double), ("beta", C.c_double)]


class GridDesc(C.Structure):
    _fields_ = [("num_blocks", C.c_int * 3), ("cells_in_block", C.c_int * 3), ("num_exch", C.c_int),
                ("lo", C.c_double * 3), ("hi", C.c_double * 3), ("periodic", C.c_int * 3),
                ("map", Map1d * 3)]


class FluxTerm(C.Structure):
    _fields_ = [("kind", C.c_int), ("order", C.c_int), ("sensor", C.c_int),
                ("sensor_param", C.c_double), ("eps_p", C.c_double), ("eps_T", C.c_double),
                ("split", C.c_int), ("smooth", C.c_int), ("law", C.c_int),
                ("mu", C.c_double), ("T_ref", C.c_double), ("power", C.c_double),
                ("prandtl", C.c_double), ("diss_kind", C.c_int), ("full", C.c_int),
                ("sgs", C.c_int), ("sgs_cw", C.c_double), ("sgs_delta", C.c_double), ("sgs_prt", C.c_double)]


class FillKernel(C.Structure):
    _fields_ = [("dom", C.c_double * 5), ("fill", C.c_double * 5), ("add", C.c_double * 5),
                ("dom_normal", C.c_double), ("fill_normal", C.c_double), ("add_normal", C.c_double)]


class FluxDesc(C.Structure):
    _fields_ = [("nterms", C.c_int), ("terms", FluxTerm * MAX_TERMS), ("gamma", C.c_double), ("R", C.c_double)]


class CaseDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("cells", C.c_int * 3), ("mach", C.c_double), ("reynolds", C.c_double),
                ("gamma", C.c_double), ("prandtl", C.c_double), ("law", C.c_int), ("power", C.c_double),
                ("cfl", C.c_double)]


CASE_TGV, CASE_VORTEX = 0, 1


class BCDesc(C.Structure):
    _fields_ = [("wall_mask", C.c_int), ("t_wall", C.c_double)]


_vp = C.c_void_p
_dp = C.POINTER(C.c_double)

RK_MAX_STAGES = 6
RK2, RK2HS, RK4, SSPRK34, SSPRK3_TABLE, SSPRK3HS, RK38R = range(7)
TRANS_PRIM_CONS, TRANS_IDENTITY = 0, 1
TO_CONS, TO_PRIM = 0, 1


class Ratio(C.Structure):
    _fields_ = [("num", C.c_long), ("den", C.c_long)]


class TuneEntry(C.Structure):
    _fields_ = [("grid", C.c_int), ("interleaved", C.c_int), ("rows", C.c_int), ("prefetch", C.c_int),
                ("occupancy", C.c_int), ("ms_per_step", C.c_float)]


class RKTable(C.Structure):
    _fields_ = [("stages", C.c_int), ("high_storage", C.c_int),
                ("a", Ratio * RK_MAX_STAGES * RK_MAX_STAGES), ("b", Ratio * RK_MAX_STAGES),
                ("c", Ratio * RK_MAX_STAGES)]


RHS_FN = C.CFUNCTYPE(C.c_int, _vp, _vp, C.c_double, _vp)
BC_FN = C.CFUNCTYPE(C.c_int, _vp, C.c_double, _vp)


class RKDesc(C.Structure):
    _fields_ = [("table", RKTable), ("trans", C.c_int), ("gamma", C.c_double), ("R", C.c_double),
                ("t0", C.c_double), ("dt", C.c_double), ("rhs", RHS_FN), ("bc", BC_FN), ("user", _vp),
                ("flux", C.POINTER(FluxDesc)), ("walls", C.POINTER(BCDesc))]

# name -> (restype, argtypes); every function declared in include/spade_amd.h
SIGNATURES = {
    "spd_last_error": (C.c_char_p, []),
    "spd_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "spd_ctx_destroy": (C.c_int, [_vp]),
    "spd_ctx_sync": (C.c_int, [_vp]),
    "spd_ctx_stream": (_vp, [_vp]),
    "spd_timer_start": (C.c_int, [_vp]),
    "spd_timer_stop": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "spd_grid_create": (C.c_int, [_vp, C.POINTER(GridDesc), C.POINTER(_vp)]),
    "spd_grid_destroy": (C.c_int, [_vp]),
    "spd_array_create": (C.c_int, [_vp, C.POINTER(_vp)]),
    "spd_array_destroy": (C.c_int, [_vp]),
    "spd_array_upload": (C.c_int, [_vp, _dp]),
    "spd_array_download": (C.c_int, [_vp, _dp]),
    "spd_array_fill": (C.c_int, [_vp, C.c_double]),
    "spd_array_copy": (C.c_int, [_vp, _vp]),
    "spd_array_host_size": (C.c_long, [_vp]),
    "spd_array_device_size": (C.c_long, [_vp]),
    "spd_flux_div": (C.c_int, [_vp, _vp, C.POINTER(FluxDesc), C.c_int]),
    "spd_flux_div_kernel": (C.c_int, [_vp, C.POINTER(FluxDesc), C.POINTER(C.c_int)]),
    "spd_exchange": (C.c_int, [_vp]),
    "spd_boundary_fill_wall": (C.c_int, [_vp, C.c_int, C.c_double]),
    "spd_boundary_fill_extrap": (C.c_int, [_vp, C.c_int, C.c_int]),
    "spd_boundary_fill": (C.c_int, [_vp, C.c_int, C.POINTER(FillKernel)]),
    "spd_integrator_create": (C.c_int, [_vp, C.POINTER(FluxDesc), C.POINTER(BCDesc), C.c_int, C.c_double,
                                        C.POINTER(_vp)]),
    "spd_integrator_fused": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "spd_integrator_advance": (C.c_int, [_vp, C.c_int]),
    "spd_integrator_destroy": (C.c_int, [_vp]),
    "spd_integrator_residual": (_vp, [_vp, C.c_int]),
    "spd_rk_table_builtin": (C.c_int, [C.c_int, C.POINTER(RKTable)]),
    "spd_rk_integrator_create": (C.c_int, [_vp, C.POINTER(RKDesc), C.POINTER(_vp)]),
    "spd_integrator_time": (C.c_int, [_vp, _dp]),
    "spd_rk_stage": (C.c_int, [_vp, C.POINTER(_vp), _dp, C.c_int, C.c_double, C.c_double, C.c_int]),
    "spd_array_scale": (C.c_int, [_vp, C.c_double]),
    "spd_array_add": (C.c_int, [_vp, _vp]),
    "spd_array_sub": (C.c_int, [_vp, _vp]),
    "spd_array_convert": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_int]),
    "spd_max_wavespeed": (C.c_int, [_vp, C.c_double, C.c_double, _dp]),
    "spd_max_wavespeed_noalloc": (C.c_int, [_vp, C.c_double, C.c_double, _vp]),
    "spd_diagnostics": (C.c_int, [_vp, C.c_double, _dp, _dp]),
    "spd_batch_create": (C.c_int, [_vp, C.c_longlong, C.c_int, C.POINTER(_vp)]),
    "spd_batch_destroy": (C.c_int, [_vp]),
    "spd_batch_capacity": (C.c_int, [_vp, C.POINTER(CaseDesc), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "spd_batch_add_case": (C.c_int, [_vp, C.POINTER(CaseDesc), C.POINTER(C.c_int)]),
    "spd_batch_remove_case": (C.c_int, [_vp, C.c_int]),
    "spd_batch_step": (C.c_int, [_vp, C.c_int]),
    "spd_batch_case_info": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_int), _dp]),
    "spd_batch_case_status": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int)]),
    "spd_batch_case_download": (C.c_int, [_vp, C.c_int, _dp]),
    "spd_batch_case_upload": (C.c_int, [_vp, C.c_int, _dp]),
    "spd_batch_case_diag": (C.c_int, [_vp, C.c_int, _dp, _dp, C.c_int, C.POINTER(C.c_int)]),
    "spd_batch_tune": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "spd_batch_use_graph": (C.c_int, [_vp, C.c_int]),
    "spd_batch_set_shape": (C.c_int, [_vp, C.c_int, C.c_int]),
    "spd_batch_set_kernel_shape": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "spd_batch_get_shape": (C.c_int, [_vp] + [C.POINTER(C.c_int)]*5),
    "spd_batch_tune_log": (C.c_int, [_vp, C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "spd_io_output_vtk": (C.c_int, [_vp, C.c_char_p, C.c_char_p, C.POINTER(C.c_char_p)]),
    "spd_io_output_vtk_host": (C.c_int, [C.POINTER(GridDesc), _dp, C.c_char_p, C.c_char_p, C.POINTER(C.c_char_p)]),
    "spd_io_binary_write": (C.c_int, [_vp, C.c_char_p]),
    "spd_io_binary_read": (C.c_int, [_vp, C.c_char_p]),
}

_lib = None


class SpadeError(RuntimeError):
    """Raised for any non-zero status of the C ABI (reference: except::sp_exception)."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def load():
    """Load libspade_amd.so from the package folder (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python spade_amd/build.py`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code):
    if code != SPD_OK:
        msg = load().spd_last_error().decode()
        raise SpadeError(code, msg)
