"""Python mirror of the reference's public solver interface over the C-ABI.

Same names, argument meaning and error behaviour as include/conicxx/solver.h,
settings.h, cone_spec.h and solution.h: ``Solver.setup`` returns False on
invalid input, ``Solver.updateData`` returns False on a pattern mismatch or a
failed factorization (pass None for unchanged pieces), ``Solver.solve`` always
returns a Solution with a definite Status and never raises for numerical
trouble.  P is the upper triangle (entries below the diagonal are ignored by
the KKT system), matrices are scipy.sparse CSC (int32 indices, float64
values), cone rows are ordered Zero, Nonnegative, SOC...
"""
import ctypes
import enum
from dataclasses import dataclass, field, fields

import numpy as np
import scipy.sparse as sp

from . import _lib


class Status(enum.IntEnum):  # include/conicxx/solution.h:9-16
    Unsolved = 0
    Solved = 1
    PrimalInfeasible = 2
    DualInfeasible = 3
    MaxIterations = 4
    NumericalError = 5


def toString(status):  # src/conicxx/solution.cpp:5-21
    return _lib.lib().cx_status_string(int(status)).decode()


class LinearSolverBackend(enum.IntEnum):  # include/conicxx/settings.h:8-17
    Eigen = 0
    Qdldl = 1
    RegularizedLdlt = 2


class FactorPrecision(enum.IntEnum):
    FP64 = 0
    FP32Refined = 1


@dataclass
class Settings:  # include/conicxx/settings.h:21-62 (defaults identical)
    tol_feas: float = 1e-8
    tol_gap_abs: float = 1e-8
    tol_gap_rel: float = 1e-8
    tol_infeas: float = 1e-7
    max_iter: int = 200
    static_reg_P: float = 1e-8
    static_reg_A: float = 1e-8
    dynamic_reg_eps: float = 1e-14
    dynamic_reg_delta: float = 1e-7
    refine_max_iter: int = 3
    refine_tol: float = 1e-12
    max_step_fraction: float = 0.99
    equilibrate: bool = True
    equilibrate_max_iter: int = 10
    equilibrate_min_scale: float = 1e-4
    equilibrate_max_scale: float = 1e4
    warm_start: bool = True
    verbose: int = 0
    record_timings: bool = False
    validate_inputs: bool = True
    linear_solver: LinearSolverBackend = LinearSolverBackend.Qdldl
    factor_precision: FactorPrecision = FactorPrecision.FP64
    fp32_reg: float = 1e-4  # FP32 factorization: diagonal shift and pivot floor

    def to_c(self):
        c = _lib.CxSettings()
        for name, ctype in _lib.CxSettings._fields_:
            v = getattr(self, name)
            setattr(c, name, int(v) if ctype is ctypes.c_int else float(v))
        return c

    @classmethod
    def from_c(cls, c):
        kw = {}
        for f in fields(cls):
            v = getattr(c, f.name)
            if f.name == "linear_solver":
                v = LinearSolverBackend(v)
            elif f.name == "factor_precision":
                v = FactorPrecision(v)
            elif f.type is bool:
                v = bool(v)
            kw[f.name] = v
        return cls(**kw)


@dataclass
class ConeSpec:  # include/conicxx/cone_spec.h:15-33
    zero_dim: int = 0
    nonneg_dim: int = 0
    soc_dims: list = field(default_factory=list)

    def totalDim(self):
        return int(self.zero_dim) + int(self.nonneg_dim) + int(sum(self.soc_dims))

    def isValid(self):
        return self.zero_dim >= 0 and self.nonneg_dim >= 0 and all(d >= 1 for d in self.soc_dims)


@dataclass
class Info:  # include/conicxx/solution.h:20-28
    iterations: int = 0
    primal_residual: float = 0.0
    dual_residual: float = 0.0
    duality_gap: float = 0.0
    mu: float = 0.0
    setup_time_s: float = 0.0
    solve_time_s: float = 0.0


@dataclass
class Solution:  # include/conicxx/solution.h:30-39
    status: Status = Status.Unsolved
    x: np.ndarray = None
    s: np.ndarray = None
    z: np.ndarray = None
    objective: float = 0.0
    info: Info = field(default_factory=Info)

    def ok(self):
        return self.status == Status.Solved


def _as_csc(M, name, upper=False):
    """Compressed columns for setup(): dense P / A and CSR matrices go through
    the library's conversions (src/conicxx/solver.cpp:17-30), which warn on
    stderr; compressed-column input is taken as it is."""
    from .convert import to_sparse, to_sparse_upper_triangular
    if not sp.issparse(M):
        C = to_sparse_upper_triangular(M, name) if upper else to_sparse(M, name)
    elif isinstance(M, (sp.csr_matrix, sp.csr_array)):
        C = to_sparse(M, name)
    else:
        C = sp.csc_matrix(M)
    C = C.copy()
    C.sort_indices()
    return C


class Solver:
    """conicxx::Solver (include/conicxx/solver.h:33-95) over one device handle."""

    def __init__(self):
        L = _lib.lib()
        h = L.cx_create()
        if not h:
            raise RuntimeError("cx_create failed: no usable GPU (the solver runs on the device only)")
        self._h = ctypes.c_void_p(h)
        self._n = self._m = 0
        self._P = self._A = None
        self._solution = Solution()
        self._settings = Settings()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.lib().cx_destroy(h)
            self._h = None

    # -- setup / update -----------------------------------------------------
    def setup(self, P, q, A, b, cone_spec, settings=None):
        settings = settings or Settings()
        if not isinstance(cone_spec, ConeSpec):
            cone_spec = ConeSpec(*cone_spec)
        Pc = _as_csc(P, "P", upper=True)
        Ac = _as_csc(A, "A")
        q = np.asarray(q, dtype=np.float64).reshape(-1)
        b = np.asarray(b, dtype=np.float64).reshape(-1)
        n = Pc.shape[0]
        m = cone_spec.totalDim()
        if (not cone_spec.isValid() or Pc.shape[0] != Pc.shape[1] or Ac.shape != (m, n)
                or q.size != n or b.size != m):
            return False
        self._n, self._m = n, m
        self._settings = settings
        self._P, self._A = Pc, Ac  # patterns (the gradients of backward() live on them)
        keep = []
        a1, p1 = _lib.i32(Pc.indptr); a2, p2 = _lib.i32(Pc.indices); a3, p3 = _lib.f64(Pc.data)
        a4, p4 = _lib.i32(Ac.indptr); a5, p5 = _lib.i32(Ac.indices); a6, p6 = _lib.f64(Ac.data)
        a7, p7 = _lib.f64(q); a8, p8 = _lib.f64(b)
        sd, sdp = _lib.i32(np.asarray(cone_spec.soc_dims, dtype=np.int32).reshape(-1))
        keep += [a1, a2, a3, a4, a5, a6, a7, a8, sd]
        st = settings.to_c()
        ok = _lib.lib().cx_setup(self._h, n, m, p1, p2, p3, p7, p4, p5, p6, p8,
                                 int(cone_spec.zero_dim), int(cone_spec.nonneg_dim), len(sd), sdp,
                                 ctypes.byref(st))
        self._solution = Solution()
        return bool(ok)

    def updateData(self, P=None, q=None, A=None, b=None):
        keep = []
        p1 = p2 = _lib.NULL_I
        p3 = _lib.NULL_F
        p4 = p5 = _lib.NULL_I
        p6 = _lib.NULL_F
        p7 = p8 = _lib.NULL_F
        if P is not None:
            if not sp.issparse(P) or P.shape != (self._n, self._n):
                return False
            Pc = sp.csc_matrix(P)
            a1, p1 = _lib.i32(Pc.indptr); a2, p2 = _lib.i32(Pc.indices); a3, p3 = _lib.f64(Pc.data)
            keep += [a1, a2, a3]
        if A is not None:
            if not sp.issparse(A) or A.shape != (self._m, self._n):
                return False
            Ac = sp.csc_matrix(A)
            a4, p4 = _lib.i32(Ac.indptr); a5, p5 = _lib.i32(Ac.indices); a6, p6 = _lib.f64(Ac.data)
            keep += [a4, a5, a6]
        if q is not None:
            q = np.asarray(q, dtype=np.float64).reshape(-1)
            if q.size != self._n:
                return False
            a7, p7 = _lib.f64(q); keep.append(a7)
        if b is not None:
            b = np.asarray(b, dtype=np.float64).reshape(-1)
            if b.size != self._m:
                return False
            a8, p8 = _lib.f64(b); keep.append(a8)
        return bool(_lib.lib().cx_update_data(self._h, p1, p2, p3, p7, p4, p5, p6, p8))

    def setWarmStart(self, x, s, z):
        a, ap = _lib.f64(x); b_, bp = _lib.f64(s); c, cp = _lib.f64(z)
        _lib.lib().cx_set_warm_start(self._h, ap, bp, cp)

    def setSettings(self, settings):
        self._settings = settings
        st = settings.to_c()
        _lib.lib().cx_set_settings(self._h, ctypes.byref(st))

    def settings(self):
        st = _lib.CxSettings()
        _lib.lib().cx_get_settings(self._h, ctypes.byref(st))
        return Settings.from_c(st)

    # -- solve ----------------------------------------------------------------
    def solve(self):
        _lib.lib().cx_solve(self._h)
        return self._fetch()

    def solution(self):
        return self._solution

    def _fetch(self):
        n, m = self._n, self._m
        x = np.zeros(n); s = np.zeros(m); z = np.zeros(m)
        obj = ctypes.c_double(0.0)
        info = _lib.CxInfo()
        st = _lib.lib().cx_get_solution(self._h, x.ctypes.data_as(_lib._f64p),
                                        s.ctypes.data_as(_lib._f64p), z.ctypes.data_as(_lib._f64p),
                                        ctypes.byref(obj), ctypes.byref(info))
        self._solution = Solution(
            status=Status(st), x=x, s=s, z=z, objective=obj.value,
            info=Info(info.iterations, info.primal_residual, info.dual_residual, info.duality_gap,
                      info.mu, info.setup_time_s, info.solve_time_s))
        return self._solution

    # -- extensions (no counterpart in the reference) ----------------------------
    def kkt_stats(self):
        out = np.zeros(8)
        _lib.lib().cx_get_kkt_stats(self._h, out.ctypes.data_as(_lib._f64p))
        keys = ["N", "nnz_K", "supernodes", "levels", "nnz_L", "flops", "max_front", "nnz_L_exact"]
        return dict(zip(keys, out.tolist()))

    def profile(self):
        out = np.zeros(13)
        _lib.lib().cx_get_profile(self._h, out.ctypes.data_as(_lib._f64p))
        keys = ["factor_ms", "solve_ms", "other_ms", "factorizations", "kkt_solves",
                "trail_avg_ms", "trail_launches", "trail_flops", "fp32_factorizations", "fp64_fallbacks",
                "max_kkt_rel_residual", "fallback_rel_residual", "fallback_iteration"]
        return dict(zip(keys, out.tolist()))

    KERNEL_CLASSES = ["trail", "panel", "assemble", "small", "fwd", "bwd"]

    def kernel_stats(self):
        """Device time per kernel class of the last solve (CONICXX_KERNEL_TIMING=N:
        HIP events around the launches of every Nth factorization and KKT solve,
        scaled to all launches).  Factorizations stopped at a failing pivot are
        left out with their launches, time and work."""
        out = np.zeros(3 * len(self.KERNEL_CLASSES))
        n = _lib.lib().cx_get_kernel_stats(self._h, out.ctypes.data_as(_lib._f64p), len(self.KERNEL_CLASSES))
        return {self.KERNEL_CLASSES[k]: dict(ms=out[3 * k], launches=int(out[3 * k + 1]), work=out[3 * k + 2])
                for k in range(n)}

    def kkt_factor_solve(self, rhs, s=None, z=None, refine_max_iter=-1):
        """KktSystem::updateScalingAndFactorize + KktSystem::solve
        (kkt_system.cpp:139-150, 296-311) on the current data: the NT scaling at
        (s, z) (None: the identity scaling), one factorization with the retry
        ladder, one solve of K x = rhs with up to refine_max_iter refinement
        steps (negative: the settings' value).  Vectors are in the solver's
        internal space (the original units with equilibrate=False).  Returns
        (x, relative residual, numeric factorizations run)."""
        n, m = self._n, self._m
        r, rp = _lib.f64(np.asarray(rhs, dtype=np.float64).reshape(-1))
        if r.size != n + m:
            raise ValueError("rhs must have n + m entries")
        keep = [r]
        sp_ = zp_ = _lib.NULL_F
        if s is not None:
            a, sp_ = _lib.f64(np.asarray(s, dtype=np.float64).reshape(-1))
            b, zp_ = _lib.f64(np.asarray(z, dtype=np.float64).reshape(-1))
            if a.size != m or b.size != m:
                raise ValueError("s and z must have m entries")
            keep += [a, b]
        x = np.zeros(n + m)
        nf = ctypes.c_int(0)
        rel = _lib.lib().cx_kkt_factor_solve(self._h, sp_, zp_, rp, x.ctypes.data_as(_lib._f64p),
                                             int(refine_max_iter), ctypes.byref(nf))
        return x, float(rel), int(nf.value)

    def backward(self, dl_dx=None, dl_ds=None, dl_dz=None):
        """Adjoint pass of the last Solved solve: gradients of a loss L given
        dL/dx, dL/ds, dL/dz (None = zero).  Returns dict(P=csc on P's pattern,
        q=ndarray, A=csc on A's pattern, b=ndarray).  One adjoint KKT solve on
        the final factorization (no counterpart in the reference)."""
        n, m = self._n, self._m
        keep = []

        def vec(v, k):
            if v is None:
                return _lib.NULL_F
            a, p = _lib.f64(np.asarray(v, dtype=np.float64).reshape(-1))
            if a.size != k:
                raise ValueError("gradient has the wrong length")
            keep.append(a)
            return p
        gP = np.zeros(self._P.nnz); gq = np.zeros(n); gA = np.zeros(self._A.nnz); gb = np.zeros(m)
        ok = _lib.lib().cx_backward(self._h, vec(dl_dx, n), vec(dl_ds, m), vec(dl_dz, m),
                                    gP.ctypes.data_as(_lib._f64p), gq.ctypes.data_as(_lib._f64p),
                                    gA.ctypes.data_as(_lib._f64p), gb.ctypes.data_as(_lib._f64p))
        if not ok:
            raise RuntimeError("backward needs a Solved solve() immediately before it")
        P = sp.csc_matrix((gP, self._P.indices.copy(), self._P.indptr.copy()), shape=self._P.shape)
        A = sp.csc_matrix((gA, self._A.indices.copy(), self._A.indptr.copy()), shape=self._A.shape)
        return {"P": P, "q": gq, "A": A, "b": gb}

    # snake_case aliases
    update_data = updateData
    set_warm_start = setWarmStart
    set_settings = setSettings


def analyze_pattern(K_lower):
    """Host symbolic analysis of a symmetric pattern given by its lower triangle."""
    K = sp.csc_matrix(K_lower)
    K.sort_indices()
    N = K.shape[0]
    a1, p1 = _lib.i32(K.indptr)
    a2, p2 = _lib.i32(K.indices)
    perm = np.zeros(N, dtype=np.int32)
    cc = np.zeros(N, dtype=np.int32)
    st = np.zeros(8)
    ok = _lib.lib().cx_analyze_pattern(N, p1, p2, perm.ctypes.data_as(_lib._i32p),
                                       cc.ctypes.data_as(_lib._i32p), st.ctypes.data_as(_lib._f64p))
    if not ok:
        raise RuntimeError("symbolic analysis failed")
    keys = ["N", "nnz_K", "supernodes", "levels", "nnz_L", "flops", "max_front", "nnz_L_exact"]
    return perm, cc, dict(zip(keys, st.tolist()))


def equilibrate(P, A, cone_spec, settings=None):
    """Host Ruiz equilibration exactly as setup() computes it: (d, e, c)."""
    P = sp.csc_matrix(P); P.sort_indices()
    A = sp.csc_matrix(A); A.sort_indices()
    n, m = P.shape[0], A.shape[0]
    cs = cone_spec if isinstance(cone_spec, ConeSpec) else ConeSpec(*cone_spec)
    keep = [_lib.i32(P.indptr), _lib.i32(P.indices), _lib.f64(P.data),
            _lib.i32(A.indptr), _lib.i32(A.indices), _lib.f64(A.data),
            _lib.i32(np.asarray(cs.soc_dims, dtype=np.int32).reshape(-1))]
    st = (settings or Settings()).to_c()
    d = np.zeros(n); e = np.zeros(m); c = ctypes.c_double(0.0)
    ok = _lib.lib().cx_equilibrate(n, m, keep[0][1], keep[1][1], keep[2][1], keep[3][1], keep[4][1],
                                   keep[5][1], cs.zero_dim, cs.nonneg_dim, len(cs.soc_dims), keep[6][1],
                                   ctypes.byref(st), d.ctypes.data_as(_lib._f64p), e.ctypes.data_as(_lib._f64p),
                                   ctypes.byref(c))
    if not ok:
        raise ValueError("invalid sizes")
    return d, e, c.value
