"""ctypes view of the CPU oracle (test infrastructure only).

Loaded only by the test-suite, the smoke entry and the CPU-baseline leg of
the benchmark, as the checker.  The product package never imports this.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)


class OcSettings(ctypes.Structure):
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
    ]


def build():
    """Compile the oracle from its C source (make in this directory)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oc_create.restype = ctypes.c_void_p
        L.oc_destroy.argtypes = [ctypes.c_void_p]
        L.oc_default_settings.argtypes = [ctypes.POINTER(OcSettings)]
        L.oc_setup.restype = ctypes.c_int
        L.oc_setup.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p,
                               _f64p, _i32p, _i32p, _f64p, _f64p, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, _i32p, ctypes.POINTER(OcSettings)]
        L.oc_update_data.restype = ctypes.c_int
        L.oc_update_data.argtypes = [ctypes.c_void_p, _i32p, _i32p, _f64p, _f64p, _i32p, _i32p,
                                     _f64p, _f64p]
        L.oc_set_warm_start.argtypes = [ctypes.c_void_p, _f64p, _f64p, _f64p]
        L.oc_set_settings.argtypes = [ctypes.c_void_p, ctypes.POINTER(OcSettings)]
        L.oc_solve.restype = ctypes.c_int
        L.oc_solve.argtypes = [ctypes.c_void_p]
        L.oc_get_solution.restype = ctypes.c_int
        L.oc_get_solution.argtypes = [ctypes.c_void_p, _f64p, _f64p, _f64p, _f64p, _i32p]
        L.oc_get_trace.restype = ctypes.c_int
        L.oc_get_trace.argtypes = [ctypes.c_void_p, _f64p, ctypes.c_int]
        L.oc_get_kkt_stats.argtypes = [ctypes.c_void_p, _f64p]
        L.oc_soc_scaling.argtypes = [ctypes.c_int, _f64p, _f64p, _f64p, _f64p]
        L.oc_soc_max_step.restype = ctypes.c_double
        L.oc_soc_max_step.argtypes = [ctypes.c_int, _f64p, _f64p, ctypes.c_double]
        L.oc_soc_product.argtypes = [ctypes.c_int, _f64p, _f64p, _f64p]
        L.oc_soc_inverse_product.argtypes = [ctypes.c_int, _f64p, _f64p, _f64p]
        L.oc_cone_margins.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _f64p,
                                      _f64p, _f64p]
        L.oc_cone_max_step.restype = ctypes.c_double
        L.oc_cone_max_step.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _f64p,
                                       _f64p, ctypes.c_double]
        L.oc_cone_scaling_apply.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p,
                                            _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]
        L.oc_equilibrate.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p, _i32p, _i32p,
                                     _f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p,
                                     ctypes.POINTER(OcSettings), _f64p, _f64p, _f64p]
        L.oc_kkt_solve.restype = ctypes.c_double
        L.oc_kkt_solve.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p, _i32p, _i32p,
                                   _f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _f64p,
                                   _f64p, ctypes.POINTER(OcSettings), _f64p, _f64p, _i32p]
        _lib = L
    return _lib


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_i32p)


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_f64p)


def settings(**overrides):
    s = OcSettings()
    lib().oc_default_settings(ctypes.byref(s))
    for k, v in overrides.items():
        setattr(s, k, v)
    return s


def _csc(M):
    M = M.tocsc()
    M.sort_indices()
    return M


class OracleSolver:
    """Object view mirroring Solver::setup/updateData/setWarmStart/solve."""

    def __init__(self):
        self._h = ctypes.c_void_p(lib().oc_create())
        self._keep = []
        self.n = self.m = 0

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oc_destroy(self._h)
            self._h = None

    def setup(self, P, q, A, b, zero_dim=0, nonneg_dim=0, soc_dims=(), **st):
        P = _csc(P)
        A = _csc(A)
        self.n, self.m = P.shape[0], A.shape[0]
        Pp, Ppp = _i(P.indptr)
        Pi, Pip = _i(P.indices)
        Px, Pxp = _f(P.data)
        Ap, App = _i(A.indptr)
        Ai, Aip = _i(A.indices)
        Ax, Axp = _f(A.data)
        qq, qp = _f(q)
        bb, bp = _f(b)
        sd, sdp = _i(np.asarray(soc_dims, dtype=np.int32).reshape(-1))
        self._st = settings(**st)
        ok = lib().oc_setup(self._h, self.n, self.m, Ppp, Pip, Pxp, qp, App, Aip, Axp, bp,
                            int(zero_dim), int(nonneg_dim), len(sd), sdp,
                            ctypes.byref(self._st))
        return bool(ok)

    def update_data(self, P=None, q=None, A=None, b=None):
        nul_i = ctypes.cast(None, _i32p)
        nul_f = ctypes.cast(None, _f64p)
        keep = []
        if P is not None:
            P = _csc(P)
            a1, p1 = _i(P.indptr); a2, p2 = _i(P.indices); a3, p3 = _f(P.data)
            keep += [a1, a2, a3]
        else:
            p1, p2, p3 = nul_i, nul_i, nul_f
        if A is not None:
            A = _csc(A)
            a4, p4 = _i(A.indptr); a5, p5 = _i(A.indices); a6, p6 = _f(A.data)
            keep += [a4, a5, a6]
        else:
            p4, p5, p6 = nul_i, nul_i, nul_f
        if q is not None:
            a7, p7 = _f(q); keep.append(a7)
        else:
            p7 = nul_f
        if b is not None:
            a8, p8 = _f(b); keep.append(a8)
        else:
            p8 = nul_f
        return bool(lib().oc_update_data(self._h, p1, p2, p3, p7, p4, p5, p6, p8))

    def set_warm_start(self, x, s, z):
        a, ap = _f(x); b_, bp = _f(s); c, cp = _f(z)
        lib().oc_set_warm_start(self._h, ap, bp, cp)

    def set_settings(self, **st):
        self._st = settings(**st)
        lib().oc_set_settings(self._h, ctypes.byref(self._st))

    def solve(self):
        status = lib().oc_solve(self._h)
        x = np.zeros(self.n); s = np.zeros(self.m); z = np.zeros(self.m)
        info = np.zeros(5); it = ctypes.c_int(0)
        lib().oc_get_solution(self._h, x.ctypes.data_as(_f64p), s.ctypes.data_as(_f64p),
                              z.ctypes.data_as(_f64p), info.ctypes.data_as(_f64p),
                              ctypes.byref(it))
        return dict(status=int(status), x=x, s=s, z=z, objective=float(info[4]),
                    primal_residual=float(info[0]), dual_residual=float(info[1]),
                    duality_gap=float(info[2]), mu=float(info[3]), iterations=int(it.value))

    def trace(self):
        out = np.zeros((1000, 8))
        r = lib().oc_get_trace(self._h, out.ctypes.data_as(_f64p), 1000)
        return out[:r].copy()

    def kkt_stats(self):
        out = np.zeros(4)
        lib().oc_get_kkt_stats(self._h, out.ctypes.data_as(_f64p))
        return dict(N=int(out[0]), nnz_K=int(out[1]), nnz_L=int(out[2]), flops=float(out[3]))


def solve(P, q, A, b, zero_dim=0, nonneg_dim=0, soc_dims=(), **st):
    s = OracleSolver()
    ok = s.setup(P, q, A, b, zero_dim, nonneg_dim, soc_dims, **st)
    if not ok:
        return None
    return s.solve()


def soc_scaling(s, z):
    d = len(s)
    W = np.zeros((d, d), order="F"); H = np.zeros((d, d), order="F")
    a, ap = _f(s); b, bp = _f(z)
    lib().oc_soc_scaling(d, ap, bp, W.ctypes.data_as(_f64p), H.ctypes.data_as(_f64p))
    return W, H


def soc_max_step(x, dx, amax):
    a, ap = _f(x); b, bp = _f(dx)
    return lib().oc_soc_max_step(len(x), ap, bp, float(amax))


def soc_product(u, v):
    out = np.zeros(len(u)); a, ap = _f(u); b, bp = _f(v)
    lib().oc_soc_product(len(u), ap, bp, out.ctypes.data_as(_f64p))
    return out


def soc_inverse_product(u, w):
    out = np.zeros(len(u)); a, ap = _f(u); b, bp = _f(w)
    lib().oc_soc_inverse_product(len(u), ap, bp, out.ctypes.data_as(_f64p))
    return out


def cone_margins(zero_dim, nonneg_dim, soc_dims, x):
    sd, sdp = _i(np.asarray(soc_dims, dtype=np.int32).reshape(-1)); a, ap = _f(x)
    mn = ctypes.c_double(); ps = ctypes.c_double()
    lib().oc_cone_margins(zero_dim, nonneg_dim, len(sd), sdp, ap, ctypes.byref(mn), ctypes.byref(ps))
    return mn.value, ps.value


def cone_max_step(zero_dim, nonneg_dim, soc_dims, x, dx, amax):
    sd, sdp = _i(np.asarray(soc_dims, dtype=np.int32).reshape(-1)); a, ap = _f(x); b, bp = _f(dx)
    return lib().oc_cone_max_step(zero_dim, nonneg_dim, len(sd), sdp, ap, bp, float(amax))


def cone_scaling_apply(zero_dim, nonneg_dim, soc_dims, s, z, v):
    sd, sdp = _i(np.asarray(soc_dims, dtype=np.int32).reshape(-1))
    a, ap = _f(s); b, bp = _f(z); c, cp = _f(v)
    m = len(s)
    outs = [np.zeros(m) for _ in range(4)]
    lib().oc_cone_scaling_apply(zero_dim, nonneg_dim, len(sd), sdp, ap, bp, cp,
                                *[o.ctypes.data_as(_f64p) for o in outs])
    return tuple(outs)  # lambda=W z, Hs v, W v, W^-1 v


def equilibrate(P, A, zero_dim, nonneg_dim, soc_dims, **st):
    P = _csc(P); A = _csc(A)
    n, m = P.shape[0], A.shape[0]
    Pp, Ppp = _i(P.indptr); Pi, Pip = _i(P.indices); Px, Pxp = _f(P.data)
    Ap, App = _i(A.indptr); Ai, Aip = _i(A.indices); Ax, Axp = _f(A.data)
    sd, sdp = _i(np.asarray(soc_dims, dtype=np.int32).reshape(-1))
    stt = settings(**st)
    d = np.zeros(n); e = np.zeros(m); c = ctypes.c_double()
    lib().oc_equilibrate(n, m, Ppp, Pip, Pxp, App, Aip, Axp, zero_dim, nonneg_dim, len(sd), sdp,
                         ctypes.byref(stt), d.ctypes.data_as(_f64p), e.ctypes.data_as(_f64p),
                         ctypes.byref(c))
    return d, e, c.value


def kkt_solve(P, A, zero_dim, nonneg_dim, soc_dims, rhs, s=None, z=None, **st):
    P = _csc(P); A = _csc(A)
    n, m = P.shape[0], A.shape[0]
    Pp, Ppp = _i(P.indptr); Pi, Pip = _i(P.indices); Px, Pxp = _f(P.data)
    Ap, App = _i(A.indptr); Ai, Aip = _i(A.indices); Ax, Axp = _f(A.data)
    sd, sdp = _i(np.asarray(soc_dims, dtype=np.int32).reshape(-1))
    stt = settings(**st)
    r, rp = _f(rhs)
    x = np.zeros(n + m)
    ok = ctypes.c_int(0)
    nul = ctypes.cast(None, _f64p)
    if s is not None:
        sa, sp = _f(s); za, zp = _f(z)
    else:
        sp = zp = nul
    rel = lib().oc_kkt_solve(n, m, Ppp, Pip, Pxp, App, Aip, Axp, zero_dim, nonneg_dim, len(sd),
                             sdp, sp, zp, ctypes.byref(stt), rp, x.ctypes.data_as(_f64p),
                             ctypes.byref(ok))
    return x, rel, bool(ok.value)


# ---------------------------------------------------------------------------
# The same oracle in x87 extended precision (liboracle_xp.so: every oc_real is
# a long double, 64-bit significand).  Used by the tests only, to compute a
# problem's optimum to far below the FP64 method's rounding-level spread.
# ---------------------------------------------------------------------------
_LIB_XP_PATH = os.path.join(_HERE, "liboracle_xp.so")
_xp = None
_xpp = ctypes.POINTER(ctypes.c_longdouble)


class OcSettingsXP(ctypes.Structure):
    _fields_ = [(name, ctypes.c_longdouble if t is ctypes.c_double else t) for name, t in OcSettings._fields_]


def lib_xp():
    global _xp
    if _xp is None:
        if not os.path.exists(_LIB_XP_PATH):
            build()
        L = ctypes.CDLL(_LIB_XP_PATH)
        L.oc_create.restype = ctypes.c_void_p
        L.oc_destroy.argtypes = [ctypes.c_void_p]
        L.oc_default_settings.argtypes = [ctypes.POINTER(OcSettingsXP)]
        L.oc_setup.restype = ctypes.c_int
        L.oc_setup.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _xpp,
                               _xpp, _i32p, _i32p, _xpp, _xpp, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, _i32p, ctypes.POINTER(OcSettingsXP)]
        L.oc_solve.restype = ctypes.c_int
        L.oc_solve.argtypes = [ctypes.c_void_p]
        L.oc_get_solution.restype = ctypes.c_int
        L.oc_get_solution.argtypes = [ctypes.c_void_p, _xpp, _xpp, _xpp, _xpp, _i32p]
        _xp = L
    return _xp


def _x(a):
    a = np.ascontiguousarray(a, dtype=np.longdouble)
    return a, a.ctypes.data_as(_xpp)


def solve_xp(P, q, A, b, zero_dim=0, nonneg_dim=0, soc_dims=(), **st):
    """The oracle's IPM in extended precision; returns the same dict as solve()
    with x, s, z as np.longdouble arrays."""
    assert np.dtype(np.longdouble).itemsize > 8, "long double is no wider than double here"
    L = lib_xp()
    P = _csc(P)
    A = _csc(A)
    n, m = P.shape[0], A.shape[0]
    keep = [_i(P.indptr), _i(P.indices), _x(P.data), _i(A.indptr), _i(A.indices), _x(A.data), _x(q), _x(b),
            _i(np.asarray(soc_dims, dtype=np.int32).reshape(-1))]
    s = OcSettingsXP()
    L.oc_default_settings(ctypes.byref(s))
    for k, v in st.items():
        setattr(s, k, v)
    h = ctypes.c_void_p(L.oc_create())
    try:
        ok = L.oc_setup(h, n, m, keep[0][1], keep[1][1], keep[2][1], keep[6][1], keep[3][1], keep[4][1],
                        keep[5][1], keep[7][1], int(zero_dim), int(nonneg_dim), len(keep[8][0]), keep[8][1],
                        ctypes.byref(s))
        if not ok:
            return None
        status = L.oc_solve(h)
        x = np.zeros(n, dtype=np.longdouble); sv = np.zeros(m, dtype=np.longdouble)
        z = np.zeros(m, dtype=np.longdouble); info = np.zeros(5, dtype=np.longdouble)
        it = ctypes.c_int(0)
        L.oc_get_solution(h, x.ctypes.data_as(_xpp), sv.ctypes.data_as(_xpp), z.ctypes.data_as(_xpp),
                          info.ctypes.data_as(_xpp), ctypes.byref(it))
    finally:
        L.oc_destroy(h)
    return dict(status=int(status), x=x, s=sv, z=z, objective=info[4], primal_residual=info[0],
                dual_residual=info[1], duality_gap=info[2], mu=info[3], iterations=int(it.value))
