"""ctypes declarations of libbolt_hip.so (include/bolt_hip.h).

The library is built in-tree by bolt.jl_amd/build.py; importing this module
fails loudly when it is missing -- there is no fallback path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# BOLT_HIP_LIB selects another in-tree build of the same library (the
# NaN-filled-workspace debug build, libbolt_hip_nanfill.so)
LIB_PATH = os.environ.get("BOLT_HIP_LIB") or os.path.join(os.path.dirname(_HERE), "libbolt_hip.so")

_D = ctypes.c_double
_I = ctypes.c_int
_I64 = ctypes.c_int64
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)
_PI64 = ctypes.POINTER(ctypes.c_int64)
_VP = ctypes.c_void_p

BOLT_OK = 0
BOLT_E_ARG, BOLT_E_DEVICE, BOLT_E_NOMEM, BOLT_E_NUMERIC, BOLT_E_STATE = -1, -2, -3, -4, -5
NTAB = 16
NSCAL = 5
(T_HC, T_HC_P, T_HC_PP, T_ETA, T_ETA_P, T_ETA_PP, T_RHO0M, T_XE, T_TAU, T_TAU_P, T_TAU_PP,
 T_G, T_G_P, T_G_PP, T_TMAT, T_CSB2) = range(16)
S_H0, S_ETA0, S_RHO_CRIT, S_OMEGA_L, S_TAU0 = range(5)


class CosmoParamsC(ctypes.Structure):
    _fields_ = [(n, _D) for n in "h Omega_r Omega_b Omega_c A n Y_p N_nu Sum_m_nu".split()]


class TablesC(ctypes.Structure):
    _fields_ = [("nx", _I), ("nq", _I), ("x0", _D), ("x1", _D), ("scal", _D * NSCAL),
                ("xq", _PD), ("wq", _PD), ("coef", _PD)]


class HierCfgC(ctypes.Structure):
    _fields_ = [("ell_g", _I), ("ell_nu", _I), ("ell_mnu", _I), ("nq", _I), ("reltol", _D),
                ("abstol", _D), ("max_steps", _I), ("flags", _I)]


class StatsC(ctypes.Structure):
    _fields_ = [("steps_accepted", _I64), ("steps_rejected", _I64), ("rhs_evals", _I64),
                ("linear_solves", _I64), ("kernel_ms", _D)]


class TuneReportC(ctypes.Structure):
    _fields_ = [("solver_sr", _I), ("solver_ms", ctypes.c_float * 4), ("los_spw", _I),
                ("los_ms", ctypes.c_float * 3)]


REQ_PK, REQ_CLTT = 0, 1
REQ_QUEUED, REQ_RUNNING, REQ_DONE = 0, 1, 2


class RequestC(ctypes.Structure):
    _fields_ = [("par", CosmoParamsC), ("kind", _I), ("nk", _I), ("kmin", _D), ("kmax", _D), ("k", _PD),
                ("ell_max", _I),
                ("nkd", _I), ("cfg", HierCfgC), ("out", _PD)]


class ResultC(ctypes.Structure):
    _fields_ = [("state", _I), ("status", _I), ("queue_ms", _D), ("host_ms", _D), ("device_ms", _D),
                ("latency_ms", _D), ("solve_ms", _D), ("joined_running", _I), ("error", ctypes.c_char * 160)]


class QueueInfoC(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_longlong), ("served", ctypes.c_longlong), ("joined_running", ctypes.c_longlong),
                ("arena_bytes", ctypes.c_longlong), ("arena_peak_bytes", ctypes.c_longlong),
                ("solver_workgroups", _I)]


# every exported symbol and its signature: (restype, argtypes)
SIGNATURES = {
    "bolt_abi_version": (_I, []),
    "bolt_last_error": (ctypes.c_char_p, []),
    "bolt_state_length": (_I, [ctypes.POINTER(HierCfgC)]),
    "bolt_host_tables": (_I, [ctypes.POINTER(CosmoParamsC), _D, _D, _I, _I, _PD, _PD, _PD, _PD]),
    "bolt_host_tables_batch": (_I, [ctypes.POINTER(CosmoParamsC), _I, _D, _D, _I, _I, _I, _PD, _PD, _PD, _PD, _PI]),
    "bolt_host_tables_jacobian": (_I, [ctypes.POINTER(CosmoParamsC), _D, _D, _I, _I, _PD, _PD, _PD, _PD, _PD, _PD]),
    "bolt_host_tables_jvp": (_I, [ctypes.POINTER(CosmoParamsC), ctypes.POINTER(CosmoParamsC), _D, _D, _I, _I, _PD, _PD,
                                  _PD, _PD, _PD, _PD]),
    "bolt_host_recfast_xe": (_I, [ctypes.POINTER(CosmoParamsC), _D, _PD, _I, _PD]),
    "bolt_host_spline_eval": (_I, [ctypes.POINTER(TablesC), _I, _I, _PD, _I, _PD]),
    "bolt_ctx_create": (_I, [_I, ctypes.POINTER(_VP)]),
    "bolt_ctx_destroy": (_I, [_VP]),
    "bolt_ctx_sync": (_I, [_VP]),
    "bolt_ws_layout": (_I, [_I, _I, _I, _I, _PI64, _PI, _PI64, _PI64, _PI64]),
    "bolt_cosmo_upload": (_I, [_VP, ctypes.POINTER(CosmoParamsC), ctypes.POINTER(TablesC), _PI]),
    "bolt_cosmo_release": (_I, [_VP, _I]),
    "bolt_cosmo_upload_jvp": (_I, [_VP, ctypes.POINTER(CosmoParamsC), ctypes.POINTER(CosmoParamsC),
                                   ctypes.POINTER(TablesC), _PD, _PD, _PI]),
    "bolt_rhs_jvp_batch": (_I, [_VP, _I, ctypes.POINTER(HierCfgC), _PD, _I, _D, _PD, _PD, _PD]),
    "bolt_initial_conditions_jvp": (_I, [_VP, _I, ctypes.POINTER(HierCfgC), _PD, _I, _PD, _PD]),
    "bolt_source_grid_jvp": (_I, [_VP, _I, ctypes.POINTER(HierCfgC), _PD, _I, _PD, _PD, _PI,
                                  ctypes.POINTER(StatsC)]),
    "bolt_source_grid_jvp_dev": (_I, [_VP, _I, ctypes.POINTER(HierCfgC), _VP, _I, _VP, _VP, _VP, _VP,
                                      ctypes.POINTER(ctypes.c_float)]),
    "bolt_boltsolve_jvp_batch": (_I, [_VP, _I, ctypes.POINTER(HierCfgC), _PD, _I, _PD, _I, _PD, _PD, _PI,
                                      ctypes.POINTER(StatsC)]),
    "bolt_rhs_batch": (_I, [_VP, _I, ctypes.POINTER(HierCfgC), _PD, _I, _D, _PD, _PD, _I]),
    "bolt_initial_conditions": (_I, [_VP, _I, ctypes.POINTER(HierCfgC), _PD, _I, _PD]),
    "bolt_boltsolve_batch": (_I, [_VP, _I, ctypes.POINTER(HierCfgC), _PD, _I, _PD, _I, _PD, _PI,
                                  ctypes.POINTER(StatsC)]),
    "bolt_boltsolve_rsa_batch": (_I, [_VP, _I, ctypes.POINTER(HierCfgC), _PD, _I, _PD, _PI,
                                      ctypes.POINTER(StatsC)]),
    "bolt_source_grid": (_I, [_VP, _I, ctypes.POINTER(HierCfgC), _PD, _I, _PD, _PI, ctypes.POINTER(StatsC)]),
    "bolt_source_grid_dev": (_I, [_VP, _I, ctypes.POINTER(HierCfgC), _VP, _I, _VP, _VP, _VP,
                                  ctypes.POINTER(ctypes.c_float)]),
    "bolt_source_grid_P": (_I, [_VP, _I, ctypes.POINTER(HierCfgC), _PD, _I, _PD, _PI, ctypes.POINTER(StatsC)]),
    "bolt_source_grid_tp": (_I, [_VP, _I, ctypes.POINTER(HierCfgC), _PD, _I, _PD, _PD, _PI,
                                 ctypes.POINTER(StatsC)]),
    "bolt_source_grid_tp_dev": (_I, [_VP, _I, ctypes.POINTER(HierCfgC), _VP, _I, _VP, _VP, _VP, _VP,
                                     ctypes.POINTER(ctypes.c_float)]),
    "bolt_source_grid_batch": (_I, [_VP, _I, _PI, ctypes.POINTER(HierCfgC), _PD, _I, _PD, _PD, _PI,
                                    ctypes.POINTER(StatsC)]),
    "bolt_source_grid_batch_dev": (_I, [_VP, _I, _PI, ctypes.POINTER(HierCfgC), _VP, _I, _VP, _VP, _VP, _VP,
                                        ctypes.POINTER(ctypes.c_float)]),
    "bolt_cltt": (_I, [_VP, _I, _PD, _PD, _I, _PD, _I, _PI, _I, _I, _PD]),
    "bolt_cltt_jvp": (_I, [_VP, _I, _PD, _PD, _PD, _I, _PD, _I, _PI, _I, _PD, _PD]),
    "bolt_cltt_jvp_dev": (_I, [_VP, _I, _VP, _VP, _VP, _I, _VP, _I, _PI, _I, _VP, _VP, ctypes.POINTER(ctypes.c_float)]),
    "bolt_cl": (_I, [_VP, _I, _PD, _PD, _PD, _I, _PD, _I, _PI, _I, _I, _PD, _PD, _PD]),
    "bolt_cl_dev": (_I, [_VP, _I, _VP, _VP, _VP, _I, _VP, _I, _PI, _I, _VP, _VP, _VP, _VP,
                         ctypes.POINTER(ctypes.c_float)]),
    "bolt_cltt_dev": (_I, [_VP, _I, _VP, _VP, _I, _VP, _I, _PI, _I, _VP, _VP, ctypes.POINTER(ctypes.c_float)]),
    "bolt_autotune": (_I, [_VP, _I, ctypes.POINTER(HierCfgC), _PD, _I, _PD, _I, _I, ctypes.POINTER(TuneReportC)]),
    "bolt_graph_spectrum_create": (_I, [_VP, _I, ctypes.POINTER(HierCfgC), _VP, _I, _VP, _I, _PI, _I, _VP, _VP, _VP,
                                        _VP, ctypes.POINTER(_VP)]),
    "bolt_graph_launch": (_I, [_VP]),
    "bolt_graph_destroy": (_I, [_VP]),
    "bolt_cltt_dev_noalloc": (_I, [_VP, _I, _VP, _VP, _I, _VP, _I, _PI, _I, _VP, _VP,
                                   ctypes.POINTER(ctypes.c_float)]),
    "bolt_plin": (_I, [_VP, _I, ctypes.POINTER(HierCfgC), _PD, _I, _D, _PD, _PI, ctypes.POINTER(StatsC)]),
    "bolt_serve_start": (_I, [_VP, _I, _D]),
    "bolt_serve_slots": (_I, [_VP]),
    "bolt_serve_start_queue": (_I, [_VP, _D, _I]),
    "bolt_serve_queue_info": (_I, [_VP, ctypes.POINTER(QueueInfoC)]),
    "bolt_submit": (_I, [_VP, ctypes.POINTER(RequestC), ctypes.POINTER(ctypes.c_uint64)]),
    "bolt_poll": (_I, [_VP, ctypes.c_uint64, ctypes.POINTER(ResultC)]),
    "bolt_wait": (_I, [_VP, ctypes.c_uint64, ctypes.POINTER(ResultC)]),
    "bolt_serve_stop": (_I, [_VP]),
    "bolt_moment_table_create": (_I, [_VP, _I, _I, _D, _D, _I, ctypes.POINTER(_VP)]),
    "bolt_moment_table_destroy": (_I, [_VP]),
    "bolt_moment_table_eval": (_I, [_VP, _PD, _I, _PD]),
    "bolt_moment_table_coefficients": (_I, [_VP, _PD]),
    "bolt_filon": (_I, [_VP, _PD, _PD, _PD, _PD, _PD, _PD, _I, _PD]),
    "bolt_filon_sum": (_I, [_VP, _PD, _PD, _PD, _PD, _I, _PD, _I, _PD]),
    "bolt_sph_moments": (_I, [_VP, _I, _PD, _PI, _PD, _I, _PD]),
    "bolt_device_tables_batch": (_I, [_VP, ctypes.POINTER(CosmoParamsC), _I, _D, _D, _I, _I, _PD, _PD, _PD, _PD,
                                      _PI]),
}

_lib = None


def lib():
    """Load libbolt_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python bolt.jl_amd/build.py` "
                              "(or __graft_entry__.build()) first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class BoltError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libbolt_hip error {code}: {msg}")
        self.code = code


def check(rc):
    if rc != BOLT_OK:
        raise BoltError(rc, lib().bolt_last_error().decode("utf-8", "replace"))
    return rc


def pd(a):
    return a.ctypes.data_as(_PD)


def pi(a):
    return a.ctypes.data_as(_PI)
