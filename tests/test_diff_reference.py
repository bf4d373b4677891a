"""Pins the dense implicit-differentiation reference (tests/diff_reference.py)
on the CPU: its gradients must match fourth-order central finite differences
of the CPU oracle's solutions, and a reference with a planted mistake (a
dropped or swapped complementarity block) must not."""
import numpy as np
import pytest
import scipy.sparse as sp

import diff_reference
import oracle_py
from conicxx_amd import problems as PR

TIGHT = dict(tol_feas=1e-10, tol_gap_abs=1e-10, tol_gap_rel=1e-10, max_iter=200, warm_start=0)
H = 1e-3


def fd_samples(p, n_each=4, seed=1):
    """(kind, index, fd) for q, b, A-value and P-value coordinates.  The IPM's
    answer is accurate to ~1e-9 here, so a 4th-order stencil with h = 1e-3
    keeps both truncation and noise near 1e-4 relative."""
    def loss(q=None, b=None, P=None, A=None):
        r = oracle_py.solve(p.P if P is None else P, p.q if q is None else q,
                            p.A if A is None else A, p.b if b is None else b,
                            p.zero_dim, p.nonneg_dim, p.soc_dims, **TIGHT)
        assert r["status"] in (1, 4)  # MaxIterations returns the last (accurate) iterate
        return p.meta["w"] @ r["x"]

    def d4(f):
        return (f(-2 * H) - 8 * f(-H) + 8 * f(H) - f(2 * H)) / (12 * H)
    rng = np.random.default_rng(seed)
    out = []
    for i in rng.choice(p.n, n_each, replace=False):
        e = np.zeros(p.n); e[i] = 1.0
        out.append(("q", i, d4(lambda t: loss(q=p.q + t * e))))
    for i in rng.choice(p.m, n_each, replace=False):
        e = np.zeros(p.m); e[i] = 1.0
        out.append(("b", i, d4(lambda t: loss(b=p.b + t * e))))
    A = sp.csc_matrix(p.A); A.sort_indices()
    for k in rng.choice(A.nnz, n_each, replace=False):
        def fa(t, k=k):
            B = A.copy(); B.data[k] += t
            return loss(A=B)
        out.append(("A", k, d4(fa)))
    P = sp.csc_matrix(p.P); P.sort_indices()
    for k in rng.choice(P.nnz, n_each, replace=False):
        def fp(t, k=k):
            B = P.copy(); B.data[k] += t
            return loss(P=B)
        out.append(("P", k, d4(fp)))
    return out


def fd_error(g, samples):
    fd = np.array([s[2] for s in samples])
    gv = np.array([g[k][i] for k, i, _ in samples])
    return np.linalg.norm(fd - gv) / np.linalg.norm(gv)


@pytest.fixture(scope="module")
def case():
    p = PR.config3_problem(0)
    r = oracle_py.solve(p.P, p.q, p.A, p.b, p.zero_dim, p.nonneg_dim, p.soc_dims, **TIGHT)
    assert r["status"] == 1
    return p, r, fd_samples(p)


def test_implicit_gradients_match_finite_differences(case):
    p, r, samples = case
    g = diff_reference.implicit_gradients(p, r["x"], r["s"], r["z"], p.meta["w"])
    assert fd_error(g, samples) < 1e-2


@pytest.mark.parametrize("plant", ["swap", "drop_soc_offdiag", "sign_b", "drop_zu_in_A"])
def test_planted_mistakes_are_caught(case, plant, monkeypatch):
    p, r, samples = case
    real = diff_reference.complementarity_blocks

    def broken(s, z, zero, nonneg, soc):
        Cs, Cz = real(s, z, zero, nonneg, soc)
        if plant == "swap":
            return Cz, Cs
        if plant == "drop_soc_offdiag":
            o = zero + nonneg
            for d in soc:
                for C in (Cs, Cz):
                    C[o:o + d, o:o + d] = np.diag(np.diag(C[o:o + d, o:o + d]))
                o += d
        return Cs, Cz
    monkeypatch.setattr(diff_reference, "complementarity_blocks", broken)
    g = diff_reference.implicit_gradients(p, r["x"], r["s"], r["z"], p.meta["w"])
    if plant == "sign_b":
        g = dict(g, b=-g["b"])
    if plant == "drop_zu_in_A":  # dL/dA_ij = -(z_i u_j + v_i x_j) without its first term
        A = sp.csc_matrix(p.A); A.sort_indices()
        v = g["b"]
        gA = np.array([-(v[A.indices[k]] * r["x"][j]) for j in range(p.n)
                       for k in range(A.indptr[j], A.indptr[j + 1])])
        g = dict(g, A=gA)
    assert fd_error(g, samples) > 5e-2


def test_implicit_gradients_at_planted_optimum_match_central_differences():
    """A strictly complementary instance with a closed-form optimum: the
    reference evaluated at the exact optimum against two-point central
    differences at step 1e-6 of the oracle solved to 1e-11 (the solution map is
    smooth there and the solves are accurate to ~1e-11, so the quotient's error
    is ~1e-5 relative at most; measured 5e-9), on 64 coordinates: 16 each of
    q, b, A and P."""
    import oracle_py
    p = PR.planted(4, n=60, zero=5, nonneg=40, soc_dims=[3, 4, 5, 6], p_kind="dense")
    w = np.random.default_rng(3).standard_normal(p.n)
    g = diff_reference.implicit_gradients(p, p.meta["x"], p.meta["s"], p.meta["z"], w)
    A = sp.csc_matrix(p.A); A.sort_indices()
    P = sp.csc_matrix(p.P); P.sort_indices()
    tight = dict(tol_feas=1e-11, tol_gap_abs=1e-11, tol_gap_rel=1e-11)
    h = 1e-6

    def loss(kind, k, t):
        d = dict(P=P, q=p.q, A=A, b=p.b)
        v = d[kind].copy()
        if kind in "qb":
            v[k] += t
        else:
            v.data[k] += t
        d[kind] = v
        r = oracle_py.solve(d["P"], d["q"], d["A"], d["b"], p.zero_dim, p.nonneg_dim, p.soc_dims, **tight)
        assert r["status"] == 1
        return w @ r["x"]

    rng = np.random.default_rng(11)
    fd, gv = [], []
    for kind in "qbAP":
        count = {"q": p.n, "b": p.m, "A": A.nnz, "P": P.nnz}[kind]
        for k in rng.choice(count, 16, replace=False):
            fd.append((loss(kind, k, h) - loss(kind, k, -h)) / (2 * h))
            gv.append(g[kind][k] if kind in "qb" else g[kind][k])
    err = np.linalg.norm(np.subtract(gv, fd)) / np.linalg.norm(fd)
    print(f"planted: central-difference relerr {err:.2e}")
    assert err < 1e-6
