"""GPU parity tests: the device solver (through the C-ABI) against the CPU oracle
on the same seeded inputs, and against the reference's known answers.

Parity bar (floating point).  The reference method amplifies rounding: the
oracle itself, rerun on inputs perturbed by 1e-15 (relative, on q), changes
its iteration count and moves x, s, z by up to ~5e-6 on these instances, and
can flip a MaxIterations stall into Solved.  Even the same method in x87
extended precision moves x by ~1e-7 between stopping tolerances 1e-10 and
1e-11: at the default 1e-8 stopping test x is only determined to ~1e-6.  So
each comparison measures that envelope (the oracle on the exact input plus 3
perturbed copies) and requires:
  * the GPU status is one the oracle reaches within the envelope;
  * x, s, z and the objective within max(1e-8, 2 * spread) relative
    (max|a-b| / max(1, max|b|)) of the nearest same-status oracle run,
    spread = the largest difference among those runs;
  * where that fails (the oracle runs cluster on one side of the tolerance
    ball while the GPU run stops elsewhere inside it), the GPU solution must
    be as close to the problem's optimum as the reference method gets: within
    2x the largest distance of the oracle's runs, or of the extended-precision
    oracle at the default tolerance (the method without FP64 rounding), from
    x*, s*, z* and the objective of the extended-precision oracle solved to
    1e-11 (1e-10 or 1e-9 where the tighter runs stall; x* is known to ~1e-7
    there), with a floor of 2e-7;
  * for Solved results, the KKT optimality certificate of the GPU solution;
  * in every case the objective within 1e-8 and x, s, z within 5e-5 of the
    nearest same-status oracle run (fixed caps above the envelope).
The 1e-8 bar on x, s, z itself is tested on instances whose optimum is known
in closed form and strictly complementary (tests/test_gpu_tight.py).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from conicxx_amd import ConeSpec, LinearSolverBackend, Settings, Solver, Status
from conicxx_amd import problems as PR

pytestmark = pytest.mark.gpu

TOL = 1e-8
NPERT = 3
# hard cap on x, s, z whatever the envelope says (measured: at most 6.8e-6 over
# every comparison of the GPU suite at the default tolerance, the oracle's own
# perturbed runs spreading by up to 1e-5)
XSZ_CAP = 5e-5


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max(initial=0.0) / max(1.0, np.abs(b).max(initial=0.0)))


def gpu_solve(p, settings=None):
    s = Solver()
    assert s.setup(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims), settings)
    return s, s.solve()


def oracle_runs(oracle, p, npert=NPERT, **st):
    """The oracle on the exact input and on npert rounding-level perturbations of q."""
    runs = [oracle.solve(p.P, p.q, p.A, p.b, p.zero_dim, p.nonneg_dim, p.soc_dims, **st)]
    rng = np.random.default_rng(12345)
    for _ in range(npert):
        q = p.q * (1.0 + 1e-15 * rng.standard_normal(p.n))
        runs.append(oracle.solve(p.P, q, p.A, p.b, p.zero_dim, p.nonneg_dim, p.soc_dims, **st))
    return runs


def _errs(sol, r):
    e = {k: rel(getattr(sol, k) if not isinstance(sol, dict) else sol[k], r[k]) for k in ("x", "s", "z")}
    obj = sol.objective if not isinstance(sol, dict) else sol["objective"]
    e["objective"] = abs(obj - r["objective"]) / max(1.0, abs(r["objective"]))
    return e


def _xp(oracle, prob, **st):
    r = oracle.solve_xp(prob.P, prob.q, prob.A, prob.b, prob.zero_dim, prob.nonneg_dim, prob.soc_dims,
                        warm_start=0, **st)
    if r is None or r["status"] != 1:
        return None
    return {k: np.asarray(r[k], dtype=np.float64) for k in ("x", "s", "z")} | {"objective": float(r["objective"])}


def optimum_xp(oracle, prob):
    """x*, s*, z*, objective of the extended-precision oracle at the tightest
    tolerance (1e-11, 1e-10, 1e-9) it converges to, with the default settings
    or with more refinement or smaller static regularization; None if none."""
    for tol in (1e-11, 1e-10, 1e-9):
        for extra in ({}, {"refine_max_iter": 10}, {"static_reg_P": 1e-12, "static_reg_A": 1e-12}):
            r = _xp(oracle, prob, tol_feas=tol, tol_gap_abs=tol, tol_gap_rel=tol, refine_tol=1e-18, **extra)
            if r is not None:
                return r | {"tol": tol}
    return None


def assert_parity(sol, runs, prob=None):
    if isinstance(runs, dict):
        runs = [runs]
    statuses = {r["status"] for r in runs}
    assert int(sol.status) in statuses, (sol.status, statuses)
    same = [r for r in runs if r["status"] == int(sol.status)]
    spread = {k: 0.0 for k in ("x", "s", "z", "objective")}
    for i in range(len(same)):
        for j in range(i + 1, len(same)):
            for k, v in _errs(same[i], same[j]).items():
                spread[k] = max(spread[k], v)
    best = min((_errs(sol, r) for r in same), key=lambda e: max(e.values()))
    print("parity", {k: f"{v:.2e}" for k, v in best.items()},
          "oracle spread", {k: f"{v:.1e}" for k, v in spread.items()},
          "iters gpu", sol.info.iterations, "oracle", [r["iterations"] for r in runs])
    # the objective at the survey's 1e-8 on every instance (measured at most
    # 6.2e-10), and x, s, z under a fixed cap
    assert best["objective"] < TOL, best
    assert max(best[k] for k in ("x", "s", "z")) < XSZ_CAP, best
    miss = [k for k, v in best.items() if v > max(TOL, 2.0 * spread[k])]
    if miss:
        # accuracy against the optimum instead of closeness to the oracle's cluster
        assert sol.status == Status.Solved and prob is not None, (miss, best, spread)
        import oracle_py
        opt = optimum_xp(oracle_py, prob)
        assert opt is not None, "the extended-precision oracle did not converge"
        e_gpu = _errs(sol, opt)
        e_orc = {k: max(_errs(r, opt)[k] for r in same) for k in e_gpu}
        # where the method itself stops at the default tolerance without FP64 rounding
        dflt = _xp(oracle_py, prob)
        e_dflt = _errs(dflt, opt) if dflt is not None else {k: 0.0 for k in e_gpu}
        print("  vs optimum (xp tol", opt["tol"], "): gpu", {k: f"{v:.1e}" for k, v in e_gpu.items()},
              "oracle runs", {k: f"{v:.1e}" for k, v in e_orc.items()},
              "xp at default tol", {k: f"{v:.1e}" for k, v in e_dflt.items()},
              "gpu vs xp at default tol", {k: f"{v:.1e}" for k, v in _errs(sol, dflt).items()} if dflt else None)
        for k in e_gpu:
            assert e_gpu[k] <= max(2e-7, 2.0 * e_orc[k], 2.0 * e_dflt[k]), (k, e_gpu[k], e_orc[k], e_dflt[k])
    if prob is not None and sol.status == Status.Solved:
        c = PR.kkt_certificate(prob, sol.x, sol.s, sol.z)
        scale = 1.0 + np.abs(prob.q).max() + np.abs(prob.b).max()
        assert c["stationarity"] < 1e-6 * scale and c["feasibility"] < 1e-6 * scale, c
        assert c["margin_s"] > -1e-7 and c["margin_z"] > -1e-7, c


@pytest.mark.parametrize("case", PR.kat_problems(), ids=lambda c: c[0])
def test_known_answers(case, oracle):
    name, p, status, x, obj, tol = case
    s, sol = gpu_solve(p)
    assert int(sol.status) == status
    if x is not None:
        np.testing.assert_allclose(sol.x, x, atol=tol)
        c = PR.kkt_certificate(p, sol.x, sol.s, sol.z)
        assert c["stationarity"] < tol and c["feasibility"] < tol and c["complementarity"] < tol
    if obj is not None:
        assert abs(sol.objective - obj) < tol
    ref = oracle.solve(p.P, p.q, p.A, p.b, p.zero_dim, p.nonneg_dim, p.soc_dims)
    assert int(sol.status) == ref["status"]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_config0_parity(seed, oracle):
    p = PR.config0(seed=seed)
    s, sol = gpu_solve(p)
    assert_parity(sol, oracle_runs(oracle, p), p)


@pytest.mark.parametrize("index", [0, 1, 5])
def test_contact_qp_parity(index, oracle):
    p = PR.config2_problem(index)
    s, sol = gpu_solve(p)
    assert_parity(sol, oracle_runs(oracle, p), p)


def test_dense_diff_qp_parity(oracle):
    p = PR.config3_problem(0)
    s, sol = gpu_solve(p)
    assert_parity(sol, oracle_runs(oracle, p), p)


@pytest.mark.parametrize("scale", [0.05, 0.1])
def test_large_sparse_scaled_parity(scale, oracle):
    """config 1 scaled down: large dense fronts (blocked panels + MFMA updates)."""
    p = PR.config1(scale=scale)
    s, sol = gpu_solve(p)
    st = s.kkt_stats()
    print("kkt", st)
    assert st["max_front"] > 256  # the blocked MFMA path is exercised
    assert_parity(sol, oracle_runs(oracle, p), p)


def test_regularized_backend(oracle):
    p = PR.config0(seed=2)
    st = Settings(linear_solver=LinearSolverBackend.RegularizedLdlt)
    s, sol = gpu_solve(p, st)
    assert_parity(sol, oracle_runs(oracle, p, linear_solver=2), p)


def test_update_reuse_and_pattern_mismatch(oracle):
    """test/solver/test_update_reuse.cpp"""
    targets = [np.array([1.0, 2.0]), np.array([-1.0, 3.0]), np.array([0.5, -0.5])]
    p0 = PR.box_qp(targets[0])
    s = Solver()
    assert s.setup(p0.P, p0.q, p0.A, p0.b, ConeSpec(0, 2, []))
    o = oracle.OracleSolver()
    assert o.setup(p0.P, p0.q, p0.A, p0.b, 0, 2, [])
    for t, c in enumerate(targets):
        p = PR.box_qp(c)
        if t:
            assert s.updateData(q=p.q)
            assert o.update_data(q=p.q)
        sol = s.solve()
        ref = o.solve()
        assert_parity(sol, ref)
        np.testing.assert_allclose(sol.x, np.maximum(c, 0), atol=1e-6)
    A_bad = sp.csc_matrix(([-1.0, 0.3, -1.0], ([0, 0, 1], [0, 1, 1])), shape=(2, 2))
    A_bad.sort_indices()
    assert not s.updateData(A=A_bad)
    assert s.setup(p0.P, p0.q, A_bad, p0.b, ConeSpec(0, 2, []))
    assert s.solve().status == Status.Solved


def test_friction_demo_warm_timesteps(oracle):
    """examples/friction_contact_demo.cpp: 4 warm-started timesteps; the oracle
    envelope runs the same sequence on rounding-level perturbations of q."""
    p0 = PR.friction_demo_problem([5.0, 1.0, 1.0])
    s = Solver()
    assert s.setup(p0.P, p0.q, p0.A, p0.b, ConeSpec(0, 0, [3]))
    orcs = [oracle.OracleSolver() for _ in range(NPERT + 1)]
    for o in orcs:
        assert o.setup(p0.P, p0.q, p0.A, p0.b, 0, 0, [3])
    rng = np.random.default_rng(7)
    for t in range(4):
        f = np.array([5.0 - t, 1.0 + 2.0 * t, 1.0 + 0.5 * t])
        assert s.updateData(q=-f)
        for i, o in enumerate(orcs):
            qi = -f * (1.0 + (1e-15 * rng.standard_normal(3) if i else 0.0))
            assert o.update_data(q=qi)
        assert_parity(s.solve(), [o.solve() for o in orcs], PR.friction_demo_problem(f))


def test_warm_start_user_point(oracle):
    p = PR.config0(seed=4)
    s = Solver()
    assert s.setup(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims))
    o = oracle.OracleSolver()
    assert o.setup(p.P, p.q, p.A, p.b, p.zero_dim, p.nonneg_dim, p.soc_dims)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(p.n); sl = np.abs(rng.standard_normal(p.m)); z = np.abs(rng.standard_normal(p.m))
    s.setWarmStart(x, sl, z)
    o.set_warm_start(x, sl, z)
    assert_parity(s.solve(), o.solve())


def test_failure_statuses(oracle):
    p = PR.box_qp([1.0, 2.0])
    s, sol = gpu_solve(p, Settings(max_iter=1))
    assert sol.status == Status.MaxIterations
    q = p.q.copy(); q[0] = np.nan
    s = Solver()
    assert s.setup(p.P, q, p.A, p.b, ConeSpec(0, 2, []))
    assert s.solve().status == Status.NumericalError
    s = Solver()
    assert s.solve().status == Status.NumericalError  # solve() before setup()
    # invalid input is rejected
    assert not s.setup(p.P, p.q, p.A, p.b, ConeSpec(0, 3, []))
    assert not s.setup(p.P, p.q, p.A, p.b, ConeSpec(-1, 3, []))


def test_no_equilibration_parity(oracle):
    p = PR.config0(seed=5)
    s, sol = gpu_solve(p, Settings(equilibrate=False))
    assert_parity(sol, oracle_runs(oracle, p, equilibrate=0), p)


def test_large_sparse_certificate_many_panel_tasks():
    """config 1 at 1/4 scale (largest front ~3.7k, ~58 row tasks per sub-panel
    launch): the oracle would need minutes, so the bar is the KKT optimality
    certificate and a converged status within the reference's iteration range."""
    p = PR.config1(scale=0.25)
    s, sol = gpu_solve(p, Settings(warm_start=False))
    st = s.kkt_stats()
    print("kkt", st, "iters", sol.info.iterations, "time", sol.info.solve_time_s)
    assert st["max_front"] > 3000
    assert sol.status == Status.Solved
    assert sol.info.iterations <= 40
    c = PR.kkt_certificate(p, sol.x, sol.s, sol.z)
    scale = 1.0 + np.abs(p.q).max() + np.abs(p.b).max()
    assert c["stationarity"] < 1e-6 * scale and c["feasibility"] < 1e-6 * scale, c
    assert c["margin_s"] > -1e-7 and c["margin_z"] > -1e-7, c


def test_config1_full_size_certificate():
    """config 1 at full size (n = 20000, m = 28500, KKT order 48500, largest
    front ~15k): status Solved within the reference's iteration range and the
    KKT optimality certificate of the returned point, recomputed on the host
    from the problem data."""
    p = PR.config1()
    s, sol = gpu_solve(p, Settings(warm_start=False))
    st = s.kkt_stats()
    print("kkt", st, "iters", sol.info.iterations, "time", sol.info.solve_time_s)
    assert st["N"] == 48500 and st["max_front"] > 10000
    assert sol.status == Status.Solved
    assert sol.info.iterations <= 40
    c = PR.kkt_certificate(p, sol.x, sol.s, sol.z)
    scale = 1.0 + np.abs(p.q).max() + np.abs(p.b).max()
    assert c["stationarity"] < 1e-6 * scale and c["feasibility"] < 1e-6 * scale, c
    assert c["margin_s"] > -1e-7 and c["margin_z"] > -1e-7, c
    assert c["complementarity"] < 1e-6 * (1.0 + abs(sol.objective)), c


def _dense_input_problem():
    # the box-constrained QP of test/solver/test_solve_dense_input.cpp:11-33:
    # min 0.5 (x0 + 1)^2 + 0.5 (x1 - 2)^2 s.t. x >= 0, optimum (0, 2)
    return np.eye(2), np.array([1.0, -2.0]), -np.eye(2), np.zeros(2), ConeSpec(0, 2, [])


def test_dense_input_matches_sparse(capfd):
    """test_solve_dense_input.cpp:37-56: dense P and A (converted by the
    library, with warnings) give the sparse input's solution."""
    P, q, A, b, cs = _dense_input_problem()
    s1 = Solver()
    assert s1.setup(sp.csc_matrix(P), q, sp.csc_matrix(A), b, cs)
    ref = s1.solve()
    assert ref.ok()
    capfd.readouterr()
    s2 = Solver()
    assert s2.setup(P, q, A, b, cs)
    err = capfd.readouterr().err
    assert "(matrix: P)" in err and "(matrix: A)" in err
    sol = s2.solve()
    assert sol.ok()
    assert np.allclose(sol.x, ref.x, atol=1e-6)


def test_mixed_dense_and_sparse_inputs():
    """test_solve_dense_input.cpp:58-78, and a row-major (CSR) A."""
    P, q, A, b, cs = _dense_input_problem()
    for Pin, Ain in ((P, sp.csc_matrix(A)), (sp.csc_matrix(P), A), (P, sp.csr_matrix(A))):
        s = Solver()
        assert s.setup(Pin, q, Ain, b, cs)
        sol = s.solve()
        assert sol.ok()
        assert np.allclose(sol.x, [0.0, 2.0], atol=1e-6), sol.x


def test_update_data_new_P_and_A_values(oracle):
    """Solver::updateData with new P and A values on the setup's pattern
    (solver_impl.cpp:57-91, kkt_system.cpp:119-137): device rescaling with the
    setup's equilibration, scatter into K, refactorization.  The updated solve
    against the oracle running the same updateData sequence, and against a
    fresh setup of the new data (same status, certificate)."""
    p = PR.config0(seed=6)
    rng = np.random.default_rng(21)
    P2 = p.P.copy()  # 1.3 P + 0.05 I: positive definite on the same pattern
    P2.data = P2.data * 1.3
    col = np.repeat(np.arange(p.n), np.diff(P2.indptr))
    P2.data[P2.indices == col] += 0.05
    A2 = p.A.copy()
    A2.data = A2.data * (1.0 + 0.2 * rng.standard_normal(A2.nnz))
    q2, b2, *_ = PR.feasible_qb(rng, P2, A2, p.zero_dim, p.nonneg_dim, p.soc_dims)
    p2 = PR.Problem(P2, q2, A2, b2, p.zero_dim, p.nonneg_dim, p.soc_dims)
    st = Settings(warm_start=False)
    s = Solver()
    assert s.setup(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims), st)
    assert s.solve().status == Status.Solved
    assert s.updateData(P=P2, q=q2, A=A2, b=b2)
    sol = s.solve()
    runs = []
    rng2 = np.random.default_rng(12345)
    for k in range(NPERT + 1):
        o = oracle.OracleSolver()
        assert o.setup(p.P, p.q, p.A, p.b, p.zero_dim, p.nonneg_dim, p.soc_dims, warm_start=0)
        assert o.solve()["status"] == 1
        qk = q2 * (1.0 + (1e-15 * rng2.standard_normal(p.n) if k else 0.0))
        assert o.update_data(P=P2, q=qk, A=A2, b=b2)
        runs.append(o.solve())
    assert_parity(sol, runs, p2)
    s2, fresh = gpu_solve(p2, st)
    assert fresh.status == sol.status == Status.Solved
    assert rel(sol.x, fresh.x) < 1e-5
