"""The device KKT factorization and solve on their own, through the C-ABI
(cx_kkt_factor_solve: KktSystem::updateScalingAndFactorize + KktSystem::solve,
src/conicxx/kkt/kkt_system.cpp:139-150, 296-311), against independent dense or
sparse FP64 solves of the same matrix, as the reference's own KKT tests do
(test/kkt/test_kkt_system.cpp:32-158, test_kkt_system_backends.cpp:36-120).

K = [P + static_reg_P I, A'; A, -(Hs + static_reg_A I)], built here from the
problem data (equilibration off, so the solver's internal space is the original
one) with the cones' Hs from the oracle's NT scaling.  Bars: the reference's
1e-8 on x and relres < 1e-10 after refinement (default settings).  The
factorization alone is checked with refinement off on the same matrices with
static_reg_A = 1e-3: with the default 1e-8 the zero-cone rows (Hs = 0) give
pivots of 1e-8 whose growth puts the unrefined backward error near 1e-8 in any
LDL^T order (the reference relies on refinement there); with 1e-3 the device
LDL^T alone must reach relres < 1e-12 (FP64).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conicxx_amd import ConeSpec, FactorPrecision, Settings, Solver
from conicxx_amd import problems as PR

pytestmark = pytest.mark.gpu

SREG = 1e-8


def hs_blocks(oracle, p, s=None, z=None):
    """Block-diagonal Hs of the cone rows: 0 on zero rows, s/z on the orthant,
    the oracle's dense NT Hs per SOC block (identity at s = z = e)."""
    blocks = [sp.csc_matrix((p.zero_dim, p.zero_dim))] if p.zero_dim else []
    off = p.zero_dim
    if p.nonneg_dim:
        d = np.ones(p.nonneg_dim) if s is None else s[off:off + p.nonneg_dim] / z[off:off + p.nonneg_dim]
        blocks.append(sp.diags(d))
        off += p.nonneg_dim
    for dk in p.soc_dims:
        if s is None:
            H = np.eye(dk)
        else:
            _, H = oracle.soc_scaling(s[off:off + dk], z[off:off + dk])
        blocks.append(sp.csc_matrix(H))
        off += dk
    return sp.block_diag(blocks, format="csc") if blocks else sp.csc_matrix((0, 0))


SREG_FACTOR = 1e-3  # static_reg_A of the unrefined factorization checks


def kkt_matrix(oracle, p, s=None, z=None, sreg_p=SREG, sreg_a=SREG):
    n, m = p.n, p.m
    H = hs_blocks(oracle, p, s, z)
    K = sp.bmat([[p.P_full() + sreg_p * sp.identity(n), p.A.T],
                 [p.A, -(H + sreg_a * sp.identity(m))]], format="csc")
    return K


def interior_point(p, seed):
    """A strictly interior (s, z) pair of the problem's cone, seeded."""
    rng = np.random.default_rng(seed)
    s, z = np.zeros(p.m), np.zeros(p.m)
    off = p.zero_dim
    s[:off] = 0.0
    z[:off] = rng.standard_normal(off)
    s[off:off + p.nonneg_dim] = rng.uniform(0.1, 2.0, p.nonneg_dim)
    z[off:off + p.nonneg_dim] = rng.uniform(0.1, 2.0, p.nonneg_dim)
    off += p.nonneg_dim
    for d in p.soc_dims:
        for v in (s, z):
            u = rng.standard_normal(d - 1)
            v[off] = np.linalg.norm(u) + rng.uniform(0.1, 1.0)
            v[off + 1:off + d] = u
        off += d
    return s, z


def setup(p, **st):
    sv = Solver()
    assert sv.setup(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims),
                    Settings(equilibrate=False, **st))
    return sv


def relres(K, x, rhs):
    return float(np.linalg.norm(rhs - K @ x) / np.linalg.norm(rhs))


def relerr(x, ref):
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


def test_reference_two_by_two_cases(oracle):
    """KktSystem.SetupAndSolveMatchesDenseReference and UpdateScalingChangesOnlyHsBlock
    (test_kkt_system.cpp:32-89): P = diag(2, 3), A = I, two nonnegative rows."""
    P = sp.csc_matrix(np.diag([2.0, 3.0]))
    A = sp.csc_matrix(np.eye(2))
    p = PR.Problem(P, np.zeros(2), A, np.zeros(2), 0, 2, [])
    sv = setup(p)
    for s, z, rhs in ((None, None, np.array([1.0, 2.0, 3.0, 4.0])),
                      (np.array([4.0, 9.0]), np.array([1.0, 1.0]), np.array([1.0, -1.0, 0.5, 2.0]))):
        Hs = np.eye(2) if s is None else np.diag(s / z)
        Kd = np.block([[np.diag([2.0, 3.0]) + SREG * np.eye(2), np.eye(2)],
                       [np.eye(2), -Hs - SREG * np.eye(2)]])
        x_ref = np.linalg.solve(Kd, rhs)
        x, rel, nf = sv.kkt_factor_solve(rhs, s, z)
        print("2x2", rel, nf, relerr(x, x_ref))
        assert rel < 1e-10 and nf == 1
        np.testing.assert_allclose(x, x_ref, rtol=0, atol=1e-8)


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("scaled", [False, True], ids=["identity", "interior"])
def test_config0_kkt_matches_dense(seed, scaled, oracle):
    p = PR.config0(seed=seed)
    s, z = interior_point(p, seed) if scaled else (None, None)
    K = kkt_matrix(oracle, p, s, z)
    rhs = np.random.default_rng(100 + seed).standard_normal(p.n + p.m)
    x_ref = np.linalg.solve(K.toarray(), rhs)
    x, rel, nf = setup(p).kkt_factor_solve(rhs, s, z)
    Kf = kkt_matrix(oracle, p, s, z, sreg_a=SREG_FACTOR)
    x0, rel0, _ = setup(p, static_reg_A=SREG_FACTOR).kkt_factor_solve(rhs, s, z, refine_max_iter=0)
    print(f"config0 seed {seed} scaled {scaled}: refined relres {rel:.2e} (host {relres(K, x, rhs):.2e}), "
          f"x err {relerr(x, x_ref):.2e}; unrefined relres {relres(Kf, x0, rhs):.2e} (device {rel0:.2e})")
    assert nf == 1
    assert rel < 1e-10 and relres(K, x, rhs) < 1e-10
    assert relerr(x, x_ref) < 1e-8
    assert relres(Kf, x0, rhs) < 1e-12 and rel0 < 1e-12
    assert relerr(x0, np.linalg.solve(Kf.toarray(), rhs)) < 1e-10


@pytest.mark.parametrize("scale", [0.05, 0.1])
def test_large_sparse_kkt_matches_sparse_lu(scale, oracle):
    """Config-1 matrices with fronts above 256 rows: blocked sub-panels, MFMA
    trailing updates and the packed write-back, against SuperLU."""
    p = PR.config1(scale=scale)
    s, z = interior_point(p, 7)
    K = kkt_matrix(oracle, p, s, z)
    rhs = np.random.default_rng(3).standard_normal(p.n + p.m)
    x_ref = spla.spsolve(K, rhs)
    sv = setup(p)
    st = sv.kkt_stats()
    assert st["max_front"] > 256
    x, rel, nf = sv.kkt_factor_solve(rhs, s, z)
    Kf = kkt_matrix(oracle, p, s, z, sreg_a=SREG_FACTOR)
    x0, rel0, _ = setup(p, static_reg_A=SREG_FACTOR).kkt_factor_solve(rhs, s, z, refine_max_iter=0)
    print(f"config1 x{scale} front {st['max_front']:.0f}: refined relres {relres(K, x, rhs):.2e}, "
          f"x err {relerr(x, x_ref):.2e}; unrefined relres {relres(Kf, x0, rhs):.2e}")
    assert rel < 1e-10 and relres(K, x, rhs) < 1e-10
    assert relerr(x, x_ref) < 1e-8
    assert relres(Kf, x0, rhs) < 1e-12 and rel0 < 1e-12


def test_front_over_32768_rows_wide_panels(oracle):
    """A dense root front above 32768 rows takes the 1024-column outer panels:
    config-1 generator at 2.1x (N = 101 850, root front 32 970).  The backward
    error of an LDL^T grows with the order of its dense fronts (about f * eps,
    4e-12 at f = 33 000), so the unrefined bar here is 1e-11."""
    p = PR.config1(scale=2.1)
    Kf = kkt_matrix(oracle, p, sreg_a=SREG_FACTOR)
    rhs = np.random.default_rng(5).standard_normal(p.n + p.m)
    sv = setup(p, static_reg_A=SREG_FACTOR)
    st = sv.kkt_stats()
    assert st["max_front"] > 32768, st
    x0, rel0, nf = sv.kkt_factor_solve(rhs, refine_max_iter=0)
    print(f"config1 x2.1 front {st['max_front']:.0f}: relres unrefined {relres(Kf, x0, rhs):.2e} (device {rel0:.2e})")
    assert nf == 1
    assert relres(Kf, x0, rhs) < 1e-11 and rel0 < 1e-11


def test_update_data_reproduces_fresh_setup(oracle):
    """KktSystem.UpdateDataReproducesFreshSetup (test_kkt_system.cpp:94-121),
    on config 0: new P and A values on the same pattern through updateData give
    the same KKT solve as a fresh setup with those values."""
    p = PR.config0(seed=0)
    rng = np.random.default_rng(11)
    P2 = p.P.copy()  # 1.3 P + 0.05 I: positive definite on the same pattern
    P2.data = P2.data * 1.3
    col = np.repeat(np.arange(p.n), np.diff(P2.indptr))
    P2.data[P2.indices == col] += 0.05
    A2 = p.A.copy()
    A2.data = A2.data * (1.0 + 0.2 * rng.standard_normal(A2.nnz))
    p2 = PR.Problem(P2, p.q, A2, p.b, p.zero_dim, p.nonneg_dim, p.soc_dims)
    s, z = interior_point(p, 1)
    rhs = rng.standard_normal(p.n + p.m)
    fresh = setup(p2)
    x_fresh, rel_f, _ = fresh.kkt_factor_solve(rhs, s, z)
    reused = setup(p)
    assert reused.updateData(P=P2, A=A2)
    x_reused, rel_r, _ = reused.kkt_factor_solve(rhs, s, z)
    x_ref = np.linalg.solve(kkt_matrix(oracle, p2, s, z).toarray(), rhs)
    print("updateData", relerr(x_reused, x_fresh), relerr(x_reused, x_ref))
    assert relerr(x_reused, x_fresh) < 1e-10
    assert relerr(x_reused, x_ref) < 1e-8


def test_rank_deficient_equality_block(oracle):
    """KktSystem.HandlesRankDeficientEqualityBlockViaRegularization
    (test_kkt_system.cpp:137-158): duplicate zero-cone rows, a finite solution
    with relres < 1e-6.  With static_reg_A = 0 the second equality pivot is an
    exact zero, so the x10 ladder (kkt_system.cpp:242-264) must bump and retry:
    two factorizations, and the same solution as the oracle's ladder."""
    P = sp.csc_matrix(np.eye(2))
    A = sp.csc_matrix(np.ones((2, 2)))
    p = PR.Problem(P, np.zeros(2), A, np.zeros(2), 2, 0, [])
    rhs = np.array([1.0, 1.0, 0.0, 0.0])
    x, rel, nf = setup(p).kkt_factor_solve(rhs)
    assert np.isfinite(x).all() and rel < 1e-6 and nf == 1
    sv = setup(p, static_reg_A=0.0)
    x, rel, nf = sv.kkt_factor_solve(rhs)
    xo, relo, ok = oracle.kkt_solve(p.P, p.A, 2, 0, [], rhs, equilibrate=0, static_reg_A=0.0)
    print("ladder", x, rel, nf, "oracle", xo, relo)
    assert ok and np.isfinite(x).all()
    assert nf == 2
    assert rel < 1e-6
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-6)


def test_escalation_ladder_on_nonfinite_solve(oracle):
    """escalateAndResolve (kkt_system.cpp:266-294): a solve that stays
    non-finite after the accepted factorization runs the whole x1000 ladder
    (10 bumped refactorizations) and reports failure."""
    p = PR.config0(seed=0)
    rhs = np.ones(p.n + p.m)
    rhs[3] = np.inf
    x, rel, nf = setup(p).kkt_factor_solve(rhs)
    _, relo, _ = oracle.kkt_solve(p.P, p.A, p.zero_dim, p.nonneg_dim, p.soc_dims, rhs, equilibrate=0)
    print("escalation", rel, nf, relo)
    assert nf == 1 + 10
    assert not np.isfinite(rel) and not np.isfinite(relo)


# The retry ladder, the escalation steps and the RegularizedLdlt cap on the
# device, each against the closed form the oracle is pinned by
# (tests/test_oracle.py): static regularization and refinement off, so the
# solution is exactly a dense solve of K with the prescribed bump.  A nonnegative
# row whose only A entry is an explicit zero (Hs = 1 at s = z = e) keeps m > 0.
NO_REG = dict(static_reg_P=0.0, static_reg_A=0.0, refine_max_iter=0)


def test_retry_ladder_first_bump_on_device():
    """factorizeWithRetry (kkt_system.cpp:242-264): an exact zero pivot fails
    the first factorization; the retry bumps by +-dynamic_reg_delta (1e-7)."""
    P = sp.csc_matrix(np.diag([1.0, 0.0]))
    P.eliminate_zeros()
    A = sp.csc_matrix(np.array([[1.0, 0.0]]))
    p = PR.Problem(P, np.zeros(2), A, np.zeros(1), 1, 0, [])
    rhs = np.array([1.0, 2.0, 3.0])
    x, rel, nf = setup(p, **NO_REG).kkt_factor_solve(rhs)
    d = 1e-7
    Kb = np.array([[1.0 + d, 0.0, 1.0], [0.0, d, 0.0], [1.0, 0.0, -d]])
    print("ladder", x, nf)
    assert nf == 2
    # 1e-8: the elimination order decides rounding here (the -1e-7 row pivot
    # eliminated first turns the other pivot into 1e7); a wrong bump moves
    # x[1] = 2e7 by a factor of 10
    np.testing.assert_allclose(x, np.linalg.solve(Kb, rhs), rtol=1e-8)


def test_escalation_steps_by_1000_on_device():
    """escalateAndResolve (kkt_system.cpp:266-294): the solve overflows on the
    accepted factorization (pivot 1e-13); bumps max(1000 x last, 1e-7): 1e-7
    still overflows, 1e-4 does not.  Three factorizations."""
    P = sp.csc_matrix(np.array([[1e-13]]))
    A = sp.csc_matrix((np.array([0.0]), (np.array([0]), np.array([0]))), shape=(1, 1))  # one explicit zero
    p = PR.Problem(P, np.zeros(1), A, np.ones(1), 0, 1, [])
    rhs = np.array([1e303, 1.0])
    x, rel, nf = setup(p, **NO_REG).kkt_factor_solve(rhs)
    print("escalation", x, nf)
    assert nf == 3 and np.isfinite(x).all()
    np.testing.assert_allclose(x, [1e303 / (1e-13 + 1e-4), 1.0 / (-1.0 - 1e-4)], rtol=1e-12)


@pytest.mark.parametrize("nzero, capped", [(1, False), (2, True)])
def test_regularized_backend_five_percent_cap_on_device(nzero, capped):
    """RegularizedLdlt (regularized_ldlt.cpp:107-139, 251-253): 21 pivots; one
    corrected to +dynamic_reg_delta (4.8%) is accepted, two (9.5%) exceed the
    5% cap and the ladder bumps the diagonal instead."""
    n = 20
    dg = np.full(n, 1e-3)
    dg[:nzero] = 0.0
    P = sp.csc_matrix(np.diag(dg))
    P.eliminate_zeros()
    A = sp.csc_matrix((np.array([0.0]), (np.array([0]), np.array([n - 1]))), shape=(1, n))  # one explicit zero
    p = PR.Problem(P, np.zeros(n), A, np.ones(1), 0, 1, [])
    rhs = np.linspace(1.0, 2.0, n + 1)
    x, rel, nf = setup(p, linear_solver=2, **NO_REG).kkt_factor_solve(rhs)
    d = 1e-7
    if capped:
        expect = rhs / np.concatenate([dg + d, [-1.0 - d]])
    else:
        expect = rhs / np.concatenate([np.where(dg == 0.0, d, dg), [-1.0]])
    print("cap", nzero, nf, x[:3], expect[:3])
    assert nf == (2 if capped else 1)
    np.testing.assert_allclose(x, expect, rtol=1e-12)


def test_fp32_factorization_with_fp64_refinement(oracle):
    """Precision form on one KKT solve: the FP32 factorization preconditions
    FP64 GMRES on the exact K, which must reach refine_tol."""
    p = PR.config1(scale=0.05)
    s, z = interior_point(p, 9)
    K = kkt_matrix(oracle, p, s, z)
    rhs = np.random.default_rng(4).standard_normal(p.n + p.m)
    x_ref = spla.spsolve(K, rhs)
    sv = setup(p, factor_precision=FactorPrecision.FP32Refined)
    x, rel, nf = sv.kkt_factor_solve(rhs, s, z)
    print(f"fp32: relres {relres(K, x, rhs):.2e} (device {rel:.2e}), x err {relerr(x, x_ref):.2e}")
    assert sv.profile()["fp32_factorizations"] >= 1
    assert rel <= 1e-12 and relres(K, x, rhs) < 1e-11
    assert relerr(x, x_ref) < 1e-8
