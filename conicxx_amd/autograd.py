"""torch.autograd.Function over the solver's C-ABI (forward solve + adjoint backward).

PyTorch is only the carrier here: values go to the library as host float64
arrays, the solve and the adjoint pass (`cx_backward`: one KKT solve on the
final factorization plus the rank-one gradient assembly) run on the GPU.

    layer = QPLayer(P_pattern, A_pattern, ConeSpec(...), Settings(...))
    x, s, z = layer(P_vals, q, A_vals, b)        # optional leading batch dimension
    loss = (w * x).sum(); loss.backward()        # grads for P_vals, q, A_vals, b

P_vals / A_vals are the values on the fixed CSC patterns (P upper triangle,
in CSC order), exactly the arrays a caller of updateData() mutates.
"""
import numpy as np
import scipy.sparse as sp
import torch

from .solver import ConeSpec, Settings, Solver, Status


class QPLayer:
    """A conic QP with fixed sparsity and cones whose values are torch tensors."""

    def __init__(self, P_pattern, A_pattern, cone_spec, settings=None):
        self.P = sp.csc_matrix(P_pattern, copy=True)
        self.P.sort_indices()
        self.A = sp.csc_matrix(A_pattern, copy=True)
        self.A.sort_indices()
        self.cone = cone_spec if isinstance(cone_spec, ConeSpec) else ConeSpec(*cone_spec)
        self.settings = settings or Settings()
        self.n, self.m = self.P.shape[0], self.A.shape[0]
        self._pool = []        # one solver per batch slot, set up on first use
        self._generation = 0   # bumped by every forward; a backward must match it

    def _solver(self, k, Pv, q, Av, b):
        P = sp.csc_matrix((Pv, self.P.indices, self.P.indptr), shape=self.P.shape)
        A = sp.csc_matrix((Av, self.A.indices, self.A.indptr), shape=self.A.shape)
        while len(self._pool) <= k:
            self._pool.append(None)
        s = self._pool[k]
        if s is None or not s.updateData(P, q, A, b):
            s = Solver()
            if not s.setup(P, q, A, b, self.cone, self.settings):
                raise RuntimeError("setup failed")
            self._pool[k] = s
        return s

    def __call__(self, P_vals, q, A_vals, b):
        return ConicSolveFunction.apply(self, P_vals, q, A_vals, b)


class ConicSolveFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, layer, P_vals, q, A_vals, b):
        batched = q.dim() == 2
        dev = q.device
        Pv = P_vals.detach().to("cpu", torch.float64).numpy()
        qv = q.detach().to("cpu", torch.float64).numpy()
        Av = A_vals.detach().to("cpu", torch.float64).numpy()
        bv = b.detach().to("cpu", torch.float64).numpy()
        B = qv.shape[0] if batched else 1

        def item(a, k, nd):
            return a[k] if a.ndim == nd else a

        xs, ss, zs = [], [], []
        layer._generation += 1
        for k in range(B):
            solver = layer._solver(k, item(Pv, k, 2), item(qv, k, 2), item(Av, k, 2), item(bv, k, 2))
            sol = solver.solve()
            if sol.status != Status.Solved:
                raise RuntimeError(f"problem {k}: {sol.status.name}")
            xs.append(sol.x); ss.append(sol.s); zs.append(sol.z)
        ctx.layer, ctx.gen, ctx.B, ctx.batched = layer, layer._generation, B, batched
        ctx.shapes = (P_vals.shape, q.shape, A_vals.shape, b.shape)
        out = [torch.from_numpy(np.stack(v) if batched else v[0]).to(dev) for v in (xs, ss, zs)]
        return tuple(out)

    @staticmethod
    def backward(ctx, gx, gs, gz):
        layer = ctx.layer
        if layer._generation != ctx.gen:
            raise RuntimeError("QPLayer was called again before backward(); the factorization "
                               "of this forward pass is gone")
        dev = gx.device if gx is not None else gz.device

        def host(g):
            return None if g is None else g.detach().to("cpu", torch.float64).numpy()
        gx, gs, gz = host(gx), host(gs), host(gz)
        gP, gq, gA, gb = [], [], [], []
        for k in range(ctx.B):
            def sel(g):
                return None if g is None else (g[k] if ctx.batched else g)
            r = layer._pool[k].backward(sel(gx), sel(gs), sel(gz))
            gP.append(r["P"].data); gq.append(r["q"]); gA.append(r["A"].data); gb.append(r["b"])

        def pack(v, shape):
            t = torch.from_numpy(np.stack(v) if ctx.batched else v[0]).to(dev)
            if t.shape != shape:  # shared (unbatched) values across a batch: sum
                t = t.sum(0)
            return t
        Ps, qs, As, bs = ctx.shapes
        return None, pack(gP, Ps), pack(gq, qs), pack(gA, As), pack(gb, bs)
