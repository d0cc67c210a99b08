"""ctypes mirror of include/mfixdem.h and the loader for libmfixdem.so.

The structs below must match the C header field for field; the test suite
checks sizeof() against the library's own view (mfixdem_params_default
fills every field).  The product path is the HIP library only: loading
fails loudly when libmfixdem.so has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

MAX_PHASES = 4
MAX_PAIRS_IN = MAX_PHASES * (MAX_PHASES + 1) // 2
UNDEFINED = 9.87654321e31
ABI_VERSION = 3

LSD, HERTZ = 0, 1
EULER, ADAMS_BASHFORTH = 0, 1
FP64, MIXED = 0, 1
DRAG_GIDASPOW, DRAG_WEN_YU = 0, 1
INTERP_GARG_2012, INTERP_NONE, INTERP_SQUARE_DPVM, INTERP_GAUSS_DPVM, INTERP_LINEAR_HAT = 0, 1, 2, 3, 4
NSEARCH_NSQUARE, NSEARCH_GRID = 1, 4
DEBUG_STORE_FORCES = 1
DEBUG_TIME_KERNELS = 2
DEBUG_REFERENCE_DEPOSITION = 4
DEBUG_NO_GRAPH = 8
TIME_STEP, TIME_REBUILD, TIME_GAS, TIME_GROUPS = 0, 1, 2, 3
ROWS_THREAD, ROWS_WAVE128, ROWS_WAVE192, ROWS_WAVE256 = 0, 1, 2, 3

ERRORS = {
    0: "OK", 1: "ARG", 2: "HIP", 3: "OOM", 4: "ROGUE", 5: "EPG",
    6: "ZERO_DISTANCE", 7: "CAPACITY", 8: "NOT_READY", 9: "UNSUPPORTED", 10: "SCRATCH",
}
ERRORS_BY_NAME = {v: k for k, v in ERRORS.items()}

_D = C.c_double
_I = C.c_int32


class Params(C.Structure):
    _fields_ = [
        ("abi_version", _I), ("coll_model", _I), ("integrator", _I),
        ("precision", _I), ("nphases", _I), ("neighbor_search_n", _I),
        ("coupled", _I), ("drag_type", _I), ("interp_scheme", _I),
        ("model_b", _I), ("max_pip", _I), ("imax", _I), ("jmax", _I),
        ("kmax", _I), ("desgridsearch_imax", _I), ("desgridsearch_jmax", _I),
        ("desgridsearch_kmax", _I), ("remove_rogue", _I),
        ("des_interp_on", _I), ("des_interp_mean_fields", _I),
        ("x_min", _D), ("y_min", _D), ("z_min", _D),
        ("xlength", _D), ("ylength", _D), ("zlength", _D),
        ("dtsolid", _D), ("dtsolid_factor", _D), ("grav", _D * 3),
        ("factor_rlm", _D), ("kn", _D), ("kn_w", _D), ("kt_fac", _D),
        ("kt_w_fac", _D), ("des_etat_fac", _D), ("des_etat_w_fac", _D),
        ("mew", _D), ("mew_w", _D),
        ("des_en_input", _D * MAX_PAIRS_IN), ("des_et_input", _D * MAX_PAIRS_IN),
        ("des_en_wall_input", _D * MAX_PHASES), ("des_et_wall_input", _D * MAX_PHASES),
        ("e_young", _D * MAX_PHASES), ("v_poisson", _D * MAX_PHASES),
        ("ew_young", _D), ("vw_poisson", _D),
        ("d_p0", _D * MAX_PHASES), ("ro_s0", _D * MAX_PHASES),
        ("max_radius", _D), ("ro_g0", _D), ("mu_g0", _D), ("des_epg_clip", _D),
        ("des_interp_width", _D), ("des_neighbor_search", _I), ("des_explicitly_coupled", _I),
        ("dx", C.POINTER(C.c_double)), ("dy", C.POINTER(C.c_double)),
        ("dz", C.POINTER(C.c_double)),
    ]

    def set_cell_sizes(self, dx=None, dy=None, dz=None):
        """Non-uniform fluid grid: DX(1..IMAX), DY(1..JMAX), DZ(1..KMAX) (None
        keeps an axis uniform).  The arrays are kept with the parameters, and
        XLENGTH ... are set to their sums."""
        keep = []
        for name, arr, n, lname in (("dx", dx, self.imax, "xlength"),
                                    ("dy", dy, self.jmax, "ylength"),
                                    ("dz", dz, self.kmax, "zlength")):
            if arr is None:
                setattr(self, name, C.POINTER(C.c_double)())
                continue
            a = np.ascontiguousarray(arr, dtype=np.float64)
            if a.shape != (n,):
                raise ValueError(f"{name} needs {n} cell sizes, got {a.shape}")
            keep.append(a)
            setattr(self, name, a.ctypes.data_as(C.POINTER(C.c_double)))
            setattr(self, lname, float(a.sum()))
        self._cell_sizes = keep


class Derived(C.Structure):
    _fields_ = [
        ("dtsolid", _D), ("tcoll", _D), ("kt", _D), ("kt_w", _D),
        ("des_etan", (_D * MAX_PHASES) * MAX_PHASES),
        ("des_etat", (_D * MAX_PHASES) * MAX_PHASES),
        ("hert_kn", (_D * MAX_PHASES) * MAX_PHASES),
        ("hert_kt", (_D * MAX_PHASES) * MAX_PHASES),
        ("des_etan_wall", _D * MAX_PHASES), ("des_etat_wall", _D * MAX_PHASES),
        ("hert_kwn", _D * MAX_PHASES), ("hert_kwt", _D * MAX_PHASES),
        ("dg_xstart", _D), ("dg_ystart", _D), ("dg_zstart", _D),
        ("dg_dxinv", _D), ("dg_dyinv", _D), ("dg_dzinv", _D),
        ("dg_imax", _I), ("dg_jmax", _I), ("dg_kmax", _I), ("n_facets", _I),
    ]


class Status(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "steps", "rebuilds", "n_particles", "n_pairs", "n_contacts",
        "rogue_deleted", "epg_errors", "zero_distance", "out_of_domain")]


BC_MASS_INFLOW, BC_MASS_OUTFLOW, BC_OUTFLOW = 1, 2, 3
PLANE_W, PLANE_E, PLANE_S, PLANE_N, PLANE_B, PLANE_T = range(6)


class DemBc(C.Structure):
    """mfixdem_dem_bc: a DEM mass inlet or outlet on a face of the box."""
    _fields_ = [
        ("type", _I), ("plane", _I),
        ("x_w", _D), ("x_e", _D), ("y_s", _D), ("y_n", _D), ("z_b", _D), ("z_t", _D),
        ("u_s", _D * MAX_PHASES), ("v_s", _D * MAX_PHASES), ("w_s", _D * MAX_PHASES),
        ("ep_s", _D * MAX_PHASES), ("massflow_s", _D * MAX_PHASES),
        ("mi_start_time", _D), ("mi_end_time", _D),
    ]


def dem_bc(kind, plane, lo, hi, vel=None, ep_s=None, massflow=None, t0=0.0, t1=1.0e30):
    """A DemBc from a region lo..hi (x, y, z), per-phase velocities
    [(u, v, w), ...], BC_EP_s and BC_MASSFLOW_s (None: UNDEFINED)."""
    b = DemBc()
    b.type, b.plane = kind, plane
    b.x_w, b.y_s, b.z_b = lo
    b.x_e, b.y_n, b.z_t = hi
    for m in range(MAX_PHASES):
        b.massflow_s[m] = UNDEFINED
    for m, v in enumerate(vel or []):
        b.u_s[m], b.v_s[m], b.w_s[m] = v
    for m, e in enumerate(ep_s or []):
        b.ep_s[m] = e
    for m, f in enumerate(massflow or []):
        b.massflow_s[m] = UNDEFINED if f is None else f
    b.mi_start_time, b.mi_end_time = t0, t1
    return b


def default_params() -> Params:
    """Reference defaults; identical to mfixdem_params_default() in C."""
    p = Params()
    p.abi_version = ABI_VERSION
    p.coll_model = LSD
    p.integrator = EULER
    p.precision = FP64
    p.nphases = 1
    p.neighbor_search_n = 25
    p.coupled = 0
    p.drag_type = DRAG_GIDASPOW
    p.interp_scheme = INTERP_GARG_2012
    p.model_b = 0
    p.max_pip = 0
    p.imax = p.jmax = p.kmax = 1
    p.remove_rogue = 1
    p.dtsolid = UNDEFINED
    p.dtsolid_factor = 50.0
    p.grav[:] = [0.0, -9.80665, 0.0]
    p.factor_rlm = float(np.float32(1.2))
    p.kt_fac = float(np.float32(0.28571428571429))
    p.kt_w_fac = float(np.float32(0.28571428571429))
    p.des_etat_fac = UNDEFINED
    p.des_etat_w_fac = UNDEFINED
    for i in range(MAX_PAIRS_IN):
        p.des_en_input[i] = UNDEFINED
        p.des_et_input[i] = UNDEFINED
    for i in range(MAX_PHASES):
        p.des_en_wall_input[i] = UNDEFINED
        p.des_et_wall_input[i] = UNDEFINED
        p.e_young[i] = UNDEFINED
        p.v_poisson[i] = UNDEFINED
        p.d_p0[i] = UNDEFINED
        p.ro_s0[i] = UNDEFINED
    p.ew_young = UNDEFINED
    p.vw_poisson = UNDEFINED
    p.des_epg_clip = UNDEFINED
    p.des_interp_width = UNDEFINED
    p.des_neighbor_search = NSEARCH_GRID
    p.ro_g0 = 1.2
    p.mu_g0 = 1.8e-5
    return p


def _here() -> str:
    return os.path.dirname(os.path.abspath(__file__))


def library_path() -> str:
    return os.path.join(os.path.dirname(_here()), "lib", "libmfixdem.so")


_lib = None


def load_library():
    """Load the HIP library.  Raises when it is missing: there is no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(
            f"libmfixdem.so not built ({path}); run __graft_entry__.build()")
    lib = C.CDLL(path)
    _declare(lib)
    _lib = lib
    return lib


_P = C.c_void_p
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
_BP = C.POINTER(C.c_int8)


def _declare(lib):
    sig = {
        "mfixdem_abi_version": ([], _I),
        "mfixdem_error_string": ([_I], C.c_char_p),
        "mfixdem_params_default": ([C.POINTER(Params)], _I),
        "mfixdem_create": ([C.POINTER(Params), C.POINTER(_P)], _I),
        "mfixdem_destroy": ([_P], _I),
        "mfixdem_get_derived": ([_P, C.POINTER(Derived)], _I),
        "mfixdem_set_particles": ([_P, C.c_int64, _DP, _DP, _DP, _DP, _DP, _IP, C.c_int64], _I),
        "mfixdem_get_particles": ([_P, _DP, _DP, _DP, _BP, C.c_int64], _I),
        "mfixdem_set_gas": ([_P, _DP, _DP, _DP, _DP], _I),
        "mfixdem_get_gas": ([_P, _DP, _DP], _I),
        "mfixdem_set_gas_pressure": ([_P, _DP], _I),
        "mfixdem_get_p_force": ([_P, _DP], _I),
        "mfixdem_init": ([_P], _I),
        "mfixdem_step": ([_P, _I], _I),
        "mfixdem_run": ([_P, _I, _I], _I),
        "mfixdem_synchronize": ([_P], _I),
        "mfixdem_calc_force_dem": ([_P], _I),
        "mfixdem_calc_drag_des": ([_P], _I),
        "mfixdem_cfnewvalues": ([_P], _I),
        "mfixdem_desgrid_pic": ([_P], _I),
        "mfixdem_neighbour": ([_P], _I),
        "mfixdem_particles_in_cell": ([_P], _I),
        "mfixdem_calc_interp_weights": ([_P], _I),
        "mfixdem_comp_mean_fields": ([_P], _I),
        "mfixdem_set_debug": ([_P, _I], _I),
        "mfixdem_get_forces": ([_P, _DP, _DP, C.c_int64], _I),
        "mfixdem_get_cells": ([_P, _IP, _IP, C.c_int64], _I),
        "mfixdem_get_pairs": ([_P, _IP, _IP, _DP, C.c_int64, C.POINTER(C.c_int64)], _I),
        "mfixdem_get_status": ([_P, C.POINTER(Status)], _I),
        "mfixdem_get_stream": ([_P, C.POINTER(_P)], _I),
        "mfixdem_get_kernel_times": ([_P, _DP, C.POINTER(C.c_int64), _I], _I),
        "mfixdem_ensemble_create": ([_I, C.POINTER(Params), C.POINTER(_P)], _I),
        "mfixdem_set_tuning": ([_P, C.c_double, _I], _I),
        "mfixdem_get_device_bytes": ([_P, C.POINTER(C.c_int64)], _I),
        "mfixdem_ensemble_set_particles": ([_P, _I, C.c_int64, _DP, _DP, _DP, _DP, _DP, _IP,
                                            C.c_int64], _I),
        "mfixdem_ensemble_get_particles": ([_P, _I, _DP, _DP, _DP, _BP, C.c_int64], _I),
        "mfixdem_reset_kernel_times": ([_P], _I),
        "mfixdem_drag_gs_gas": ([_P], _I),
        "mfixdem_set_pairs": ([_P, _IP, _IP, _DP, C.c_int64], _I),
        "mfixdem_get_gas_drag": ([_P, _DP, _DP, _DP], _I),
        "mfixdem_drag_gs_gas_noalloc": ([_P, _P, _P, _P], _I),
        "mfixdem_des_time_init": ([_P], _I),
        "mfixdem_set_particle_state": ([_P, _BP, C.c_int64], _I),
        "mfixdem_set_facets": ([_P, _I, _DP, _DP, _I], _I),
        "mfixdem_set_wall_motion": ([_P, _I, _I, _DP, _DP, _DP], _I),
        "mfixdem_set_dem_bcs": ([_P, _I, C.POINTER(DemBc), C.c_uint64], _I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res


def dptr(a: np.ndarray | None):
    if a is None:
        return C.cast(None, _DP)
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], a.dtype
    return a.ctypes.data_as(_DP)


def iptr(a: np.ndarray | None):
    if a is None:
        return C.cast(None, _IP)
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_IP)


def bptr(a: np.ndarray | None):
    if a is None:
        return C.cast(None, _BP)
    assert a.dtype == np.int8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_BP)


class MfixDemError(RuntimeError):
    """A non-zero status; .code is the MFIXDEM_ERR_* value."""

    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


def check(rc: int, what: str = ""):
    if rc != 0:
        raise MfixDemError(f"{what}: status {rc} ({ERRORS.get(rc, '?')})", rc)
