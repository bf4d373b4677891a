"""Allocation-free updateData (cx_update_data_noalloc) against cx_update_data,
both called through the C-ABI: the same data through either entry point must
leave the solver in the same state, so the following solve() gives the same
status, iteration count and bit-identical x, s, z and objective.  Two calls in
a row on two problems, the second smaller than the first (the per-GPU scratch
array keeps the first call's contents), with and without equilibration, plus
the pattern check and the size limit."""
import numpy as np
import pytest
import scipy.sparse as sp

from conicxx_amd import ConeSpec, Settings, Solver, Status, _lib
from conicxx_amd import problems as PR

pytestmark = pytest.mark.gpu

SCRATCH_DOUBLES = 1 << 20


def call(name, solver, P=None, q=None, A=None, b=None):
    keep = []
    args = []
    for M in (P,):
        if M is None:
            args += [_lib.NULL_I, _lib.NULL_I, _lib.NULL_F]
        else:
            M = sp.csc_matrix(M)
            a1, p1 = _lib.i32(M.indptr); a2, p2 = _lib.i32(M.indices); a3, p3 = _lib.f64(M.data)
            keep += [a1, a2, a3]
            args += [p1, p2, p3]
    if q is None:
        args.append(_lib.NULL_F)
    else:
        a, pq = _lib.f64(np.asarray(q, dtype=np.float64)); keep.append(a); args.append(pq)
    if A is None:
        args += [_lib.NULL_I, _lib.NULL_I, _lib.NULL_F]
    else:
        M = sp.csc_matrix(A)
        a1, p1 = _lib.i32(M.indptr); a2, p2 = _lib.i32(M.indices); a3, p3 = _lib.f64(M.data)
        keep += [a1, a2, a3]
        args += [p1, p2, p3]
    if b is None:
        args.append(_lib.NULL_F)
    else:
        a, pb = _lib.f64(np.asarray(b, dtype=np.float64)); keep.append(a); args.append(pb)
    return getattr(_lib.lib(), name)(solver._h, *args)


def new_values(p, seed):
    """Same patterns, new P and A values, and q, b from a fresh feasible pair."""
    rng = np.random.default_rng(seed)
    P2 = p.P.copy()  # 1.3 P + 0.05 I stays positive definite on the same pattern (P has its diagonal)
    P2.data = P2.data * 1.3
    col = np.repeat(np.arange(p.n), np.diff(P2.indptr))
    P2.data[P2.indices == col] += 0.05
    A2 = p.A.copy()
    A2.data = A2.data * (1.0 + 0.2 * rng.standard_normal(A2.nnz))
    q2, b2, *_ = PR.feasible_qb(rng, P2, A2, p.zero_dim, p.nonneg_dim, p.soc_dims)
    return P2, q2, A2, b2


def solve_after(entry, p, upd, settings):
    s = Solver()
    assert s.setup(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims), settings)
    first = s.solve()
    assert first.status == Status.Solved
    assert call(entry, s, *upd) == 1
    return s, s.solve()


def assert_same(a, b):
    assert a.status == b.status and a.info.iterations == b.info.iterations, (a.status, b.status)
    for k in ("x", "s", "z"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert a.objective == b.objective


@pytest.mark.parametrize("equilibrate", [True, False])
def test_noalloc_update_matches_update_data(equilibrate):
    st = Settings(equilibrate=equilibrate, warm_start=False)
    big = PR.config0(seed=3)
    small = PR.config0(seed=1, n=60, zero=6, nonneg=30, n_soc=6)
    for p, seed in ((big, 5), (small, 6)):  # the second call is smaller than the first
        P2, q2, A2, b2 = new_values(p, seed)
        upd = (P2, q2, A2, b2)
        _, ref = solve_after("cx_update_data", p, upd, st)
        s, got = solve_after("cx_update_data_noalloc", p, upd, st)
        print(p.name, ref.status.name, ref.info.iterations, got.status.name, got.info.iterations)
        assert ref.status == Status.Solved
        assert_same(got, ref)
        # q / b only, on the same handle (no refactorization)
        q3 = q2 * 1.01
        assert call("cx_update_data_noalloc", s, q=q3) == 1
        got3 = s.solve()
        s_ref, _ = solve_after("cx_update_data", p, upd, st)
        assert call("cx_update_data", s_ref, q=q3) == 1
        assert_same(got3, s_ref.solve())


def test_noalloc_update_rejects_pattern_and_size():
    p = PR.config0(seed=2, n=60, zero=6, nonneg=30, n_soc=6)
    s = Solver()
    assert s.setup(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims), Settings(warm_start=False))
    before = s.solve()
    A_bad = p.A.tolil()
    row0 = set(sp.csr_matrix(p.A)[0].indices.tolist())
    A_bad[0, min(c for c in range(p.n) if c not in row0)] = 1.0
    assert call("cx_update_data_noalloc", s, A=sp.csc_matrix(A_bad)) == 0
    # a problem whose A has more entries than the scratch array holds
    rng = np.random.default_rng(0)
    n, m = 1050, 1000
    A = sp.csc_matrix(rng.standard_normal((m, n)))
    assert A.nnz > SCRATCH_DOUBLES
    P = sp.identity(n, format="csc")
    q, b, *_ = PR.feasible_qb(rng, P, A, 0, m, [])
    sb = Solver()
    assert sb.setup(P, q, A, b, ConeSpec(0, m, []), Settings(warm_start=False))
    r1 = sb.solve()
    A2 = A.copy()
    A2.data = A2.data * 1.1
    assert call("cx_update_data_noalloc", sb, A=A2) == 0
    assert call("cx_update_data_noalloc", sb, q=q * 1.0) == 1  # small pieces still go through
    r2 = sb.solve()
    assert r1.status == r2.status == Status.Solved
    np.testing.assert_allclose(r2.x, r1.x, rtol=0, atol=1e-12)
    after = s.solve()
    assert_same(after, before)
