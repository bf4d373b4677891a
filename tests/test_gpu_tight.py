"""Tight-tolerance parity tier: every GPU form at the survey's 1e-8 bar.

The instances (problems.planted_suite) have an optimum fixed in closed form
before q and b are formed, with every cone pair strictly complementary and P
positive definite, so the solution map is Lipschitz at the optimum and a
stopping tolerance of 1e-11 determines x, s and z to ~1e-11 (the oracle lands
within 1e-13..3.5e-10 of the planted point, pinned on the CPU in
tests/test_oracle.py).  Here each GPU form, solved to 1e-11, must reach:
  * status Solved, as the oracle does;
  * x, s, z within 1e-9 relative of the planted optimum (max|a-b| / max(1, max|b|));
  * x, s, z and the objective within 1e-8 relative of the oracle's solution
    under the same settings.
The differentiation form is checked against the dense implicit reference at
the planted optimum and against two-point central differences at step 1e-6 on
64 coordinates per problem.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from conicxx_amd import (BatchSolver, ConeSpec, FactorPrecision, LinearSolverBackend, Settings, Solver,
                         Status)
from conicxx_amd import problems as PR
from diff_reference import implicit_gradients

pytestmark = pytest.mark.gpu

TIGHT = dict(tol_feas=1e-11, tol_gap_abs=1e-11, tol_gap_rel=1e-11)
SUITE = PR.planted_suite()


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max(initial=0.0) / max(1.0, np.abs(b).max(initial=0.0)))


def oracle_tight(oracle, p, **st):
    r = oracle.solve(p.P, p.q, p.A, p.b, p.zero_dim, p.nonneg_dim, p.soc_dims, **TIGHT, **st)
    assert r["status"] == 1
    return r


def check_tight(sol, p, ref, planted_bar=1e-9, oracle_bar=1e-8):
    assert sol.status == Status.Solved, sol.status
    errs = {k: (rel(getattr(sol, k), p.meta[k]), rel(getattr(sol, k), ref[k])) for k in ("x", "s", "z")}
    dobj = abs(sol.objective - ref["objective"]) / max(1.0, abs(ref["objective"]))
    print(p.name, "iters", sol.info.iterations, "oracle", ref["iterations"],
          {k: f"planted {a:.1e} oracle {b:.1e}" for k, (a, b) in errs.items()}, f"objective {dobj:.1e}")
    for k, (a, b) in errs.items():
        assert a < planted_bar, (k, a)
        assert b < oracle_bar, (k, b)
    assert dobj < oracle_bar


def solve(p, **st):
    s = Solver()
    assert s.setup(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims), Settings(**TIGHT, **st))
    return s, s.solve()


@pytest.mark.parametrize("p", SUITE, ids=lambda p: p.name)
def test_plain_fp64_tight(p, oracle):
    s, sol = solve(p)
    check_tight(sol, p, oracle_tight(oracle, p))


@pytest.mark.parametrize("p", SUITE[:3] + SUITE[5:], ids=lambda p: p.name)
def test_regularized_backend_tight(p, oracle):
    s, sol = solve(p, linear_solver=LinearSolverBackend.RegularizedLdlt)
    check_tight(sol, p, oracle_tight(oracle, p, linear_solver=2))


@pytest.mark.parametrize("p", SUITE[:3], ids=lambda p: p.name)
def test_no_equilibration_tight(p, oracle):
    s, sol = solve(p, equilibrate=False)
    check_tight(sol, p, oracle_tight(oracle, p, equilibrate=0))


@pytest.mark.parametrize("p", SUITE, ids=lambda p: p.name)
def test_fp32_factorization_tight(p, oracle):
    """FP32 factor as the preconditioner of FP64 GMRES against the exact K: the
    same solution as the FP64 method at 1e-11."""
    s, sol = solve(p, factor_precision=FactorPrecision.FP32Refined)
    pr = s.profile()
    print("profile", {k: pr[k] for k in ("factorizations", "fp32_factorizations", "fp64_fallbacks")})
    assert pr["fp32_factorizations"] >= 1
    check_tight(sol, p, oracle_tight(oracle, p))


def test_batch_tight(oracle):
    """The batched persistent kernel on the planted instances that fit one
    workgroup's problem size (all but the 1000-variable one), in one launch."""
    probs = SUITE[:5]
    b = BatchSolver()
    assert b.setup([(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims)) for p in probs],
                   Settings(**TIGHT))
    st = b.solve()
    for k, p in enumerate(probs):
        assert int(st[k]) == int(Status.Solved)
        check_tight(b.solution(k), p, oracle_tight(oracle, p))


def _fd_problem():
    return PR.planted(4, n=60, zero=5, nonneg=40, soc_dims=[3, 4, 5, 6], p_kind="dense")


def test_backward_at_planted_optimum():
    """cx_backward after a 1e-11 solve against the dense implicit reference at
    the exact planted optimum (nondegenerate: the NT-scaled adjoint and the
    exact complementarity Jacobian agree to ~mu there), and against two-point
    central differences of the GPU solution map at step 1e-6 on 64 coordinates
    (16 each of q, b, A and P)."""
    p = _fd_problem()
    w = np.random.default_rng(3).standard_normal(p.n)
    s, sol = solve(p, warm_start=False)
    assert sol.status == Status.Solved
    g = s.backward(w)
    ref = implicit_gradients(p, p.meta["x"], p.meta["s"], p.meta["z"], w)
    for k in ("P", "q", "A", "b"):
        gk = g[k].data if k in "PA" else g[k]
        e = np.linalg.norm(gk - ref[k]) / np.linalg.norm(ref[k])
        print(k, f"implicit {e:.1e}")
        assert e < 1e-6, (k, e)
    A = sp.csc_matrix(p.A); A.sort_indices()
    P = sp.csc_matrix(p.P); P.sort_indices()
    cs = ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims)
    h = 1e-6

    def loss(kind, k, t):
        d = dict(P=P, q=p.q, A=A, b=p.b)
        v = d[kind].copy()
        if kind in "qb":
            v[k] += t
        else:
            v.data[k] += t
        d[kind] = v
        r = Solver()
        assert r.setup(d["P"], d["q"], d["A"], d["b"], cs, Settings(**TIGHT, warm_start=False))
        res = r.solve()
        assert res.status == Status.Solved
        return w @ res.x

    rng = np.random.default_rng(11)
    fd, gv = [], []
    for kind in "qbAP":
        count = {"q": p.n, "b": p.m, "A": A.nnz, "P": P.nnz}[kind]
        for k in rng.choice(count, 16, replace=False):
            fd.append((loss(kind, k, h) - loss(kind, k, -h)) / (2 * h))
            gv.append(g[kind][k] if kind in "qb" else g[kind].data[k])
    err = np.linalg.norm(np.subtract(gv, fd)) / np.linalg.norm(fd)
    print(f"central differences over {len(fd)} coordinates: relerr {err:.1e}")
    assert err < 1e-5


def test_batched_backward_at_planted_optimum():
    """cx_batch_backward on four planted instances against the dense implicit
    reference at each exact optimum."""
    probs = [_fd_problem()] + [PR.planted(s, n=60, zero=5, nonneg=40, soc_dims=[3, 4, 5, 6], p_kind=k)
                               for s, k in ((5, "dense"), (6, "sparse"), (8, "sparse"))]
    rng = np.random.default_rng(3)
    ws = [rng.standard_normal(p.n) for p in probs]
    b = BatchSolver()
    assert b.setup([(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims)) for p in probs],
                   Settings(**TIGHT, warm_start=False))
    b.set_differentiable(True)
    st = b.solve()
    grads = b.backward(ws)
    for k, p in enumerate(probs):
        assert int(st[k]) == int(Status.Solved)
        ref = implicit_gradients(p, p.meta["x"], p.meta["s"], p.meta["z"], ws[k])
        for key in ("P", "q", "A", "b"):
            e = np.linalg.norm(grads[k][key] - ref[key]) / np.linalg.norm(ref[key])
            print(p.name, key, f"{e:.1e}")
            assert e < 1e-6, (p.name, key, e)


@pytest.mark.parametrize("p", SUITE, ids=lambda p: p.name)
def test_default_tolerance_matches_oracle(p, oracle):
    """At the default 1e-8 stopping tolerances (where x is determined only to
    the tolerance, ~1e-8..4e-7 from the planted point) the GPU run follows the
    oracle's trajectory: same iteration count, x, s, z and objective within
    1e-8 of the oracle's, plain and batched."""
    s, sol = Solver(), None
    assert s.setup(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims))
    sol = s.solve()
    ref = oracle.solve(p.P, p.q, p.A, p.b, p.zero_dim, p.nonneg_dim, p.soc_dims)
    assert ref["status"] == 1 and sol.status == Status.Solved
    assert sol.info.iterations == ref["iterations"]
    check_tight(sol, p, ref, planted_bar=1e-6)
    if p.n <= 400:
        b = BatchSolver()
        assert b.setup([(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims))])
        b.solve()
        check_tight(b.solution(0), p, ref, planted_bar=1e-6)
