"""ctypes binding of libconicxx_hip.so (declared in include/conicxx_c.h).

The library is built in-tree by conicxx_amd/build.py.  Loading fails loudly
when it is missing: there is no CPU fallback for the solver.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libconicxx_hip.so")

_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)


class CxSettings(ctypes.Structure):
    _fields_ = [
        ("tol_feas", ctypes.c_double),
        ("tol_gap_abs", ctypes.c_double),
        ("tol_gap_rel", ctypes.c_double),
        ("tol_infeas", ctypes.c_double),
        ("max_iter", ctypes.c_int),
        ("static_reg_P", ctypes.c_double),
        ("static_reg_A", ctypes.c_double),
        ("dynamic_reg_eps", ctypes.c_double),
        ("dynamic_reg_delta", ctypes.c_double),
        ("refine_max_iter", ctypes.c_int),
        ("refine_tol", ctypes.c_double),
        ("max_step_fraction", ctypes.c_double),
        ("equilibrate", ctypes.c_int),
        ("equilibrate_max_iter", ctypes.c_int),
        ("equilibrate_min_scale", ctypes.c_double),
        ("equilibrate_max_scale", ctypes.c_double),
        ("warm_start", ctypes.c_int),
        ("verbose", ctypes.c_int),
        ("record_timings", ctypes.c_int),
        ("validate_inputs", ctypes.c_int),
        ("linear_solver", ctypes.c_int),
        ("factor_precision", ctypes.c_int),
        ("fp32_reg", ctypes.c_double),
    ]


class CxInfo(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("primal_residual", ctypes.c_double),
        ("dual_residual", ctypes.c_double),
        ("duality_gap", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("setup_time_s", ctypes.c_double),
        ("solve_time_s", ctypes.c_double),
    ]


# (name, restype, argtypes) for every symbol the header declares
SIGNATURES = [
    ("cx_default_settings", None, [ctypes.POINTER(CxSettings)]),
    ("cx_create", ctypes.c_void_p, []),
    ("cx_destroy", None, [ctypes.c_void_p]),
    ("cx_setup", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p,
                                _f64p, _i32p, _i32p, _f64p, _f64p, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, _i32p, ctypes.POINTER(CxSettings)]),
    ("cx_update_data", ctypes.c_int, [ctypes.c_void_p, _i32p, _i32p, _f64p, _f64p, _i32p, _i32p,
                                      _f64p, _f64p]),
    ("cx_update_data_noalloc", ctypes.c_int, [ctypes.c_void_p, _i32p, _i32p, _f64p, _f64p, _i32p, _i32p,
                                              _f64p, _f64p]),
    ("cx_set_warm_start", None, [ctypes.c_void_p, _f64p, _f64p, _f64p]),
    ("cx_set_settings", None, [ctypes.c_void_p, ctypes.POINTER(CxSettings)]),
    ("cx_get_settings", None, [ctypes.c_void_p, ctypes.POINTER(CxSettings)]),
    ("cx_solve", ctypes.c_int, [ctypes.c_void_p]),
    ("cx_get_solution", ctypes.c_int, [ctypes.c_void_p, _f64p, _f64p, _f64p, _f64p,
                                       ctypes.POINTER(CxInfo)]),
    ("cx_status_string", ctypes.c_char_p, [ctypes.c_int]),
    ("cx_get_dims", ctypes.c_int, [ctypes.c_void_p, _i32p, _i32p]),
    ("cx_get_kkt_stats", None, [ctypes.c_void_p, _f64p]),
    ("cx_get_profile", None, [ctypes.c_void_p, _f64p]),
    ("cx_get_kernel_stats", ctypes.c_int, [ctypes.c_void_p, _f64p, ctypes.c_int]),
    ("cx_backward", ctypes.c_int, [ctypes.c_void_p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]),
    ("cx_kkt_factor_solve", ctypes.c_double, [ctypes.c_void_p, _f64p, _f64p, _f64p, _f64p, ctypes.c_int, _i32p]),
    ("cx_batch_create", ctypes.c_void_p, []),
    ("cx_batch_destroy", None, [ctypes.c_void_p]),
    ("cx_batch_setup", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _i32p, _i32p, ctypes.POINTER(_i32p),
                                      ctypes.POINTER(_i32p), ctypes.POINTER(_f64p), ctypes.POINTER(_f64p),
                                      ctypes.POINTER(_i32p), ctypes.POINTER(_i32p), ctypes.POINTER(_f64p),
                                      ctypes.POINTER(_f64p), _i32p, _i32p, _i32p, ctypes.POINTER(_i32p),
                                      ctypes.POINTER(CxSettings)]),
    ("cx_batch_update_data", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_i32p), ctypes.POINTER(_i32p),
                                            ctypes.POINTER(_f64p), ctypes.POINTER(_f64p), ctypes.POINTER(_i32p),
                                            ctypes.POINTER(_i32p), ctypes.POINTER(_f64p), ctypes.POINTER(_f64p)]),
    ("cx_batch_solve", ctypes.c_int, [ctypes.c_void_p, _i32p]),
    ("cx_batch_get_solution", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _f64p, _f64p, _f64p, _f64p,
                                             ctypes.POINTER(CxInfo)]),
    ("cx_batch_set_settings", None, [ctypes.c_void_p, ctypes.POINTER(CxSettings)]),
    ("cx_batch_autotune", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("cx_batch_set_tuning", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ("cx_batch_get_tuning", None, [ctypes.c_void_p, _f64p]),
    ("cx_batch_get_stats", None, [ctypes.c_void_p, ctypes.c_int, _f64p]),
    ("cx_batch_set_differentiable", None, [ctypes.c_void_p, ctypes.c_int]),
    ("cx_batch_backward", ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(_f64p)] * 7),
    ("cx_batch_last_backward_ms", ctypes.c_double, [ctypes.c_void_p]),
    ("cx_batch_last_solve_ms", ctypes.c_double, [ctypes.c_void_p]),
    ("cx_equilibrate", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p,
                                      ctypes.POINTER(CxSettings), _f64p, _f64p, _f64p]),
    ("cx_analyze_pattern", ctypes.c_int, [ctypes.c_int, _i32p, _i32p, _i32p, _i32p, _f64p]),
    ("cx_dense_to_csc", ctypes.c_long, [ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_char_p,
                                        ctypes.c_int, _i32p, _i32p, _f64p]),
    ("cx_dense_to_csc_upper", ctypes.c_long, [ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                              _i32p, _i32p, _f64p]),
    ("cx_csr_to_csc", ctypes.c_long, [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p, ctypes.c_char_p,
                                      ctypes.c_int, _i32p, _i32p, _f64p]),
]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(the solver has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def i32(a):
    import numpy as np

    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_i32p)


def f64(a):
    import numpy as np

    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_f64p)


NULL_I = ctypes.cast(None, _i32p)
NULL_F = ctypes.cast(None, _f64p)
