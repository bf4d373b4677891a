"""Batched form over the C-ABI (cx_batch_*): many independent conic QPs in one
persistent launch per solve, replayed as a HIP graph across update_data()
calls.  Each problem follows the same method as conicxx_amd.Solver."""
import ctypes

import numpy as np

from . import _lib
from .solver import ConeSpec, Info, Settings, Solution, Status, _as_csc


class BatchSolver:
    def __init__(self):
        h = _lib.lib().cx_batch_create()
        if not h:
            raise RuntimeError("cx_batch_create failed: no usable GPU (the solver runs on the device only)")
        self._h = ctypes.c_void_p(h)
        self._dims = []

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.lib().cx_batch_destroy(h)
            self._h = None

    def set_differentiable(self, on=True):
        """Make every Solved problem's forward pass also factor its adjoint matrix."""
        _lib.lib().cx_batch_set_differentiable(self._h, 1 if on else 0)

    def backward(self, dl_dx=None, dl_ds=None, dl_dz=None):
        """Gradients for every problem (lists of per-problem arrays, None = zero).
        Returns a list of dict(P=values on P's pattern, q, A=values on A's pattern, b)."""
        K = len(self._dims)
        keep = []

        def inp(vals, idx):
            if vals is None:
                return None
            out = (_lib._f64p * K)()
            for k, v in enumerate(vals):
                if v is None:
                    out[k] = _lib.NULL_F
                    continue
                a, p = _lib.f64(np.asarray(v, dtype=np.float64).reshape(-1))
                if a.size != self._dims[k][idx]:
                    raise ValueError(f"problem {k}: wrong length")
                keep.append(a)
                out[k] = p
            return out
        res = [dict(P=np.zeros(self._nnz[k][0]), q=np.zeros(self._dims[k][0]), A=np.zeros(self._nnz[k][1]),
                    b=np.zeros(self._dims[k][1])) for k in range(K)]

        def outp(key):
            arr = (_lib._f64p * K)()
            for k in range(K):
                arr[k] = res[k][key].ctypes.data_as(_lib._f64p) if res[k][key].size else _lib.NULL_F
            return arr
        r = _lib.lib().cx_batch_backward(self._h, inp(dl_dx, 0), inp(dl_ds, 1), inp(dl_dz, 1), outp("P"),
                                         outp("q"), outp("A"), outp("b"))
        if r < 0:
            raise RuntimeError("batch backward failed (differentiation off, or device failure)")
        return res

    def last_backward_ms(self):
        return float(_lib.lib().cx_batch_last_backward_ms(self._h))

    def setup(self, problems, settings=None):
        """problems: sequence of (P, q, A, b, ConeSpec)."""
        settings = settings or Settings()
        K = len(problems)
        keep = []
        arrs = {k: [] for k in ("Pp", "Pi", "Px", "q", "Ap", "Ai", "Ax", "b", "sd")}
        n, m, zd, nd, ns = (np.zeros(K, np.int32) for _ in range(5))
        self._dims = []
        self._nnz = []
        for k, (P, q, A, b, cone) in enumerate(problems):
            if not isinstance(cone, ConeSpec):
                cone = ConeSpec(*cone)
            Pc = _as_csc(P, "P", upper=True)
            Ac = _as_csc(A, "A")
            q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
            b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1)
            n[k], m[k] = Pc.shape[0], Ac.shape[0]
            zd[k], nd[k], ns[k] = cone.zero_dim, cone.nonneg_dim, len(cone.soc_dims)
            self._dims.append((int(n[k]), int(m[k])))
            self._nnz.append((int(Pc.nnz), int(Ac.nnz)))
            for key, a, conv in (("Pp", Pc.indptr, _lib.i32), ("Pi", Pc.indices, _lib.i32),
                                 ("Px", Pc.data, _lib.f64), ("q", q, _lib.f64),
                                 ("Ap", Ac.indptr, _lib.i32), ("Ai", Ac.indices, _lib.i32),
                                 ("Ax", Ac.data, _lib.f64), ("b", b, _lib.f64),
                                 ("sd", np.asarray(cone.soc_dims, np.int32).reshape(-1), _lib.i32)):
                arr, ptr = conv(a if len(a) else np.zeros(1))
                keep.append(arr)
                arrs[key].append(ptr)

        def parr(lst, t):
            return (t * K)(*lst)
        ok = _lib.lib().cx_batch_setup(
            self._h, K, n.ctypes.data_as(_lib._i32p), m.ctypes.data_as(_lib._i32p),
            parr(arrs["Pp"], _lib._i32p), parr(arrs["Pi"], _lib._i32p), parr(arrs["Px"], _lib._f64p),
            parr(arrs["q"], _lib._f64p), parr(arrs["Ap"], _lib._i32p), parr(arrs["Ai"], _lib._i32p),
            parr(arrs["Ax"], _lib._f64p), parr(arrs["b"], _lib._f64p), zd.ctypes.data_as(_lib._i32p),
            nd.ctypes.data_as(_lib._i32p), ns.ctypes.data_as(_lib._i32p), parr(arrs["sd"], _lib._i32p),
            ctypes.byref(settings.to_c()))
        return bool(ok)

    def update_data(self, qs=None, bs=None, Ps=None, As=None):
        """Solver.updateData for every problem at once: lists of new q, b, P
        (upper triangle) and A per problem, None (or a None entry) = unchanged.
        P and A must keep setup's sparsity pattern; returns False, changing
        nothing, if any of them does not."""
        K = len(self._dims)
        keep = []

        def parr(vals, idx):
            if vals is None:
                return None
            out = (_lib._f64p * K)()
            for k, v in enumerate(vals):
                if v is None:
                    out[k] = _lib.NULL_F
                    continue
                a, p = _lib.f64(np.asarray(v, dtype=np.float64).reshape(-1))
                if a.size != self._dims[k][idx]:
                    raise ValueError(f"problem {k}: wrong length")
                keep.append(a)
                out[k] = p
            return out

        def marr(mats, name, upper):
            if mats is None:
                return None, None, None
            pp, ii, vv = (_lib._i32p * K)(), (_lib._i32p * K)(), (_lib._f64p * K)()
            for k, M in enumerate(mats):
                if M is None:
                    pp[k], ii[k], vv[k] = _lib.NULL_I, _lib.NULL_I, _lib.NULL_F
                    continue
                C = _as_csc(M, name, upper=upper)
                for dst, a, conv in ((pp, C.indptr, _lib.i32), (ii, C.indices, _lib.i32), (vv, C.data, _lib.f64)):
                    arr, ptr = conv(a if len(a) else np.zeros(1))
                    keep.append(arr)
                    dst[k] = ptr
            return pp, ii, vv
        Pp, Pi, Px = marr(Ps, "P", True)
        Ap, Ai, Ax = marr(As, "A", False)
        return bool(_lib.lib().cx_batch_update_data(self._h, Pp, Pi, Px, parr(qs, 0), Ap, Ai, Ax, parr(bs, 1)))

    def solve(self):
        """Solves every problem; returns the array of statuses."""
        st = np.zeros(len(self._dims), np.int32)
        r = _lib.lib().cx_batch_solve(self._h, st.ctypes.data_as(_lib._i32p))
        if r < 0:
            raise RuntimeError("batched solve failed on the device")
        return st

    def solution(self, k):
        n, m = self._dims[k]
        x = np.zeros(n); s = np.zeros(m); z = np.zeros(m)
        obj = ctypes.c_double(0.0)
        info = _lib.CxInfo()
        st = _lib.lib().cx_batch_get_solution(self._h, k, x.ctypes.data_as(_lib._f64p),
                                              s.ctypes.data_as(_lib._f64p), z.ctypes.data_as(_lib._f64p),
                                              ctypes.byref(obj), ctypes.byref(info))
        return Solution(status=Status(st), x=x, s=s, z=z, objective=obj.value,
                        info=Info(info.iterations, info.primal_residual, info.dual_residual,
                                  info.duality_gap, info.mu, info.setup_time_s, info.solve_time_s))

    def set_settings(self, settings):
        _lib.lib().cx_batch_set_settings(self._h, ctypes.byref(settings.to_c()))

    def autotune(self, reps=1):
        return int(_lib.lib().cx_batch_autotune(self._h, reps))

    def set_tuning(self, lds_doubles, grid):
        """Pins a launch configuration: lds_doubles of LDS per workgroup (what
        does not fit lives in global memory) and grid persistent workgroups."""
        return bool(_lib.lib().cx_batch_set_tuning(self._h, int(lds_doubles), int(grid)))

    def tuning(self):
        out = np.zeros(4)
        _lib.lib().cx_batch_get_tuning(self._h, out.ctypes.data_as(_lib._f64p))
        return dict(lds_doubles=int(out[0]), grid=int(out[1]), max_packed=int(out[2]), autotuned=bool(out[3]))

    def stats(self, k):
        """Iterations, factorizations, reduced KKT solves (incl. refinement) and the
        device time (us) of problem k in the last solve."""
        out = np.zeros(13)
        _lib.lib().cx_batch_get_stats(self._h, k, out.ctypes.data_as(_lib._f64p))
        return dict(iterations=int(out[0]), factorizations=int(out[1]), solves=int(out[2]), us=float(out[3]),
                    us_factor=float(out[4]), us_solve=float(out[5]), us_residual=float(out[6]),
                    kkt_solves=int(out[7]), us_backward=float(out[8]), us_backward_solve=float(out[9]),
                    backward_solves=int(out[10]), us_last_kkt_solve=float(out[11]),
                    us_adjoint_factor=float(out[12]))

    def last_solve_ms(self):
        return float(_lib.lib().cx_batch_last_solve_ms(self._h))
