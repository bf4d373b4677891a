"""Differentiation form on the GPU: cx_backward (adjoint KKT solve on the
factorization of the tangent-corrected KKT matrix at the optimum) against the
dense FP64 implicit-differentiation reference and against finite differences,
and the torch.autograd.Function wrapper over it.  Problems: the differentiation
workload (n = 64, m = 112: 16 zero, 64 nonnegative, 8 SOC of dim 4, dense P
and A, loss w'x)."""
import numpy as np
import pytest
import scipy.sparse as sp

from conicxx_amd import ConeSpec, Settings, Solver, Status
from conicxx_amd import problems as PR
from diff_reference import implicit_gradients

pytestmark = pytest.mark.gpu

TIGHT = Settings(tol_feas=1e-9, tol_gap_abs=1e-9, tol_gap_rel=1e-9, max_iter=200, warm_start=False)


def solved(p, settings=TIGHT):
    s = Solver()
    assert s.setup(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims), settings)
    sol = s.solve()
    return s, sol


def relerr(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


# measured deviation of the adjoint (tangent-corrected NT scaling) from the
# exact linearization of complementarity at the computed iterate, at the 1e-9
# stopping tolerances: 4e-5, 2.7e-4 and 7e-6 on these problems; the bars keep
# a factor of about 4 to 7 above that
IMPLICIT_BAR = {0: 2e-4, 4: 1e-3, 5: 5e-5}


@pytest.mark.parametrize("idx", sorted(IMPLICIT_BAR))
def test_backward_matches_dense_implicit_reference(idx):
    p = PR.config3_problem(idx)
    s, sol = solved(p)
    assert sol.status == Status.Solved, sol.status
    g = s.backward(p.meta["w"])
    ref = implicit_gradients(p, sol.x, sol.s, sol.z, p.meta["w"])
    errs = {k: relerr(g[k].data if k in "PA" else g[k], ref[k]) for k in ("P", "q", "A", "b")}
    print(idx, {k: f"{v:.2e}" for k, v in errs.items()})
    for k, e in errs.items():
        assert e < IMPLICIT_BAR[idx], (k, e)


@pytest.mark.parametrize("idx", [0, 1, 2, 3, 4, 5, 6, 7])
def test_backward_matches_finite_differences(idx):
    """Central finite differences on 64 random coordinates of the problem (16
    each of q, b, A's and P's stored entries), fourth-order stencil, step 1e-3.

    The step: the solution of these problems is determined to about 1e-7
    relative at the 1e-9 stopping tolerances, so a central quotient at step
    1e-6 is noise-dominated (measured relative errors 0.1 to 0.9 on these
    eight problems, all perturbed solves Solved); at step 1e-3 the fourth-order
    stencil's truncation error stays below that noise floor (measured 2e-4 to
    2.5e-3).  A coordinate counts when its four perturbed solves end Solved."""
    p = PR.config3_problem(idx)
    s, sol = solved(p)
    assert sol.status == Status.Solved
    g = s.backward(p.meta["w"])
    h = 1e-3
    rng = np.random.default_rng(100 + idx)
    w = p.meta["w"]
    cs = ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims)
    A = sp.csc_matrix(p.A); A.sort_indices()
    P = sp.csc_matrix(p.P); P.sort_indices()

    def loss(kind, k, t):
        d = dict(P=P, q=p.q, A=A, b=p.b)
        if kind in "qb":
            v = d[kind].copy(); v[k] += t
        else:
            v = d[kind].copy(); v.data[k] += t
        d[kind] = v
        r = Solver()
        assert r.setup(d["P"], d["q"], d["A"], d["b"], cs, TIGHT)
        res = r.solve()
        return w @ res.x if res.status == Status.Solved else None

    fd, gv = [], []
    for kind in "qbAP":
        count = {"q": p.n, "b": p.m, "A": A.nnz, "P": P.nnz}[kind]
        for k in rng.choice(count, 16, replace=False):
            v = [loss(kind, k, t) for t in (-2 * h, -h, h, 2 * h)]
            if any(x is None for x in v):
                continue
            fd.append((v[0] - 8 * v[1] + 8 * v[2] - v[3]) / (12 * h))
            gv.append(g[kind][k] if kind in "qb" else g[kind].data[k])
    assert len(fd) >= 56, f"only {len(fd)} of 64 coordinates had four solved perturbations"
    err = relerr(gv, fd)
    print(f"problem {idx}: fd relerr {err:.2e} over {len(fd)} coordinates")
    assert err < 1e-2


def test_backward_reuse_and_invalidation():
    p = PR.config3_problem(4)
    s, sol = solved(p)
    assert sol.status == Status.Solved
    w = p.meta["w"]
    g1 = s.backward(w)
    g2 = s.backward(w)  # second call reuses the adjoint factorization
    assert relerr(g2["q"], g1["q"]) < 1e-12
    # linearity in the loss vector
    g3 = s.backward(2.0 * w)
    assert relerr(g3["b"], 2.0 * g1["b"]) < 1e-10
    # dL/dz and dL/ds paths against the reference
    wz = np.random.default_rng(0).standard_normal(p.m)
    ws = np.random.default_rng(1).standard_normal(p.m)
    g4 = s.backward(None, ws, wz)
    ref = implicit_gradients(p, sol.x, sol.s, sol.z, np.zeros(p.n), ws, wz)
    assert relerr(g4["q"], ref["q"]) < 5e-3
    assert relerr(g4["b"], ref["b"]) < 5e-3
    # new data invalidates the factorization of the solved problem
    assert s.updateData(q=p.q * 1.01)
    with pytest.raises(RuntimeError):
        s.backward(w)


def test_autograd_wrapper_batch():
    torch = pytest.importorskip("torch")
    from conicxx_amd.autograd import QPLayer
    probs = [PR.config3_problem(i) for i in (0, 4, 5)]
    p0 = probs[0]
    layer = QPLayer(p0.P, p0.A, ConeSpec(p0.zero_dim, p0.nonneg_dim, p0.soc_dims), TIGHT)
    Pv = torch.tensor(np.stack([sp.csc_matrix(p.P).data for p in probs]), requires_grad=True)
    Av = torch.tensor(np.stack([sp.csc_matrix(p.A).data for p in probs]), requires_grad=True)
    q = torch.tensor(np.stack([p.q for p in probs]), requires_grad=True)
    b = torch.tensor(np.stack([p.b for p in probs]), requires_grad=True)
    W = torch.tensor(np.stack([p.meta["w"] for p in probs]))
    x, s, z = layer(Pv, q, Av, b)
    (W * x).sum().backward()
    for k, p in enumerate(probs):
        ref = implicit_gradients(p, x[k].detach().numpy(), s[k].detach().numpy(), z[k].detach().numpy(),
                                 p.meta["w"])
        assert relerr(q.grad[k].numpy(), ref["q"]) < 5e-3
        assert relerr(b.grad[k].numpy(), ref["b"]) < 5e-3
        assert relerr(Av.grad[k].numpy(), ref["A"]) < 5e-3
        assert relerr(Pv.grad[k].numpy(), ref["P"]) < 5e-3
    # a second call with new values goes through updateData(P, q, A, b) on the
    # pool's solvers (same patterns); its gradients against the reference at
    # the new solutions
    rng = np.random.default_rng(8)
    Pv2 = (Pv.detach() * 1.2).requires_grad_(True)
    Av2 = (Av.detach() * (1.0 + 0.02 * torch.tensor(rng.standard_normal(Av.shape)))).requires_grad_(True)
    q2 = (q.detach() + 0.01 * torch.tensor(rng.standard_normal(q.shape))).requires_grad_(True)
    b2 = (b.detach() + 0.01 * torch.tensor(np.abs(rng.standard_normal(b.shape)))).requires_grad_(True)
    x2, s2, z2 = layer(Pv2, q2, Av2, b2)
    (W * x2).sum().backward()
    assert relerr(x2.detach().numpy(), x.detach().numpy()) > 1e-4  # a different problem
    for k, p in enumerate(probs):
        pk = PR.Problem(sp.csc_matrix((Pv2[k].detach().numpy(), p.P.indices, p.P.indptr), shape=p.P.shape),
                        q2[k].detach().numpy(),
                        sp.csc_matrix((Av2[k].detach().numpy(), p.A.indices, p.A.indptr), shape=p.A.shape),
                        b2[k].detach().numpy(), p.zero_dim, p.nonneg_dim, p.soc_dims)
        ref = implicit_gradients(pk, x2[k].detach().numpy(), s2[k].detach().numpy(), z2[k].detach().numpy(),
                                 p.meta["w"])
        assert relerr(q2.grad[k].numpy(), ref["q"]) < 5e-3
        assert relerr(b2.grad[k].numpy(), ref["b"]) < 5e-3
        assert relerr(Av2.grad[k].numpy(), ref["A"]) < 5e-3
        assert relerr(Pv2.grad[k].numpy(), ref["P"]) < 5e-3


def test_batched_backward_matches_reference_and_single_problem_path():
    """cx_batch_backward over a batch of differentiation problems: each problem's
    gradients against the dense implicit-differentiation reference (at the
    batch's own solution) and against the single-problem cx_backward."""
    from conicxx_amd import BatchSolver
    idx = [0, 4, 5, 7]
    probs = [PR.config3_problem(i) for i in idx]
    b = BatchSolver()
    assert b.setup([(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims)) for p in probs], TIGHT)
    b.set_differentiable(True)
    st = b.solve()
    grads = b.backward([p.meta["w"] for p in probs])
    print("forward ms", b.last_solve_ms(), "backward ms", b.last_backward_ms())
    for k, p in enumerate(probs):
        sol = b.solution(k)
        if sol.status != Status.Solved:
            assert not grads[k]["q"].any()
            continue
        ref = implicit_gradients(p, sol.x, sol.s, sol.z, p.meta["w"])
        errs = {key: relerr(grads[k][key], ref[key]) for key in ("P", "q", "A", "b")}
        print(idx[k], {key: f"{v:.2e}" for key, v in errs.items()})
        for key, e in errs.items():
            assert e < 1e-3, (idx[k], key, e)
        # the single-problem path against the same reference at its own
        # solution: both paths are pinned to one reference, so the two agree
        # to the sum of their errors (on problem 7 each path carries ~1e-3 of
        # its own, the gradient there being ill-conditioned, so a direct bar
        # between them would compare those two errors with each other)
        s1, sol1 = solved(p)
        if sol1.status == Status.Solved:
            g1 = s1.backward(p.meta["w"])
            ref1 = implicit_gradients(p, sol1.x, sol1.s, sol1.z, p.meta["w"])
            e1 = max(relerr(g1[key], ref1[key]) for key in ("q", "b"))
            print("  single path vs its reference", f"{e1:.2e}", "batch vs single",
                  f"{max(relerr(grads[k][key], g1[key]) for key in ('q', 'b')):.2e}")
            assert e1 < 2e-3, (idx[k], e1)
    assert (st == int(Status.Solved)).sum() >= 3
