"""Deterministic synthetic conic QP instances.

Every instance is  min 1/2 x'Px + q'x  s.t.  Ax + s = b,  s in K, with K the
product Zero x Nonneg x SOC x ... x SOC in that fixed order (the reference's
ConeSpec layout, include/conicxx/cone_spec.h:9-33).  P is returned as its
upper triangle in CSC form, A in CSC form, both with sorted row indices.

Random numbers come from numpy's PCG64 seeded by a SeedSequence of
(config id, problem index, timestep), so every instance is reproducible on
any host.  N(mu, sigma) denotes a normal distribution with *standard
deviation* sigma throughout.

Feasible construction (configs 0, 1, 3, 4 and the SOC rows of config 2),
the same scheme as the reference's benchmark generators
(benchmarks/problem_generators.cpp:58-75) with the distributions of the
workload definitions: x0 ~ N(0,1)^n; s0 is 0 on zero rows, U[0.1,1] on
nonnegative rows and (||u|| + U[0.1,1], u), u ~ N(0,1)^(d-1), on each SOC
block; z0 is N(0,1) on zero rows and an interior point of the same kind on
cone rows; b = A x0 + s0 and q = -P x0 - A' z0.
"""
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp


@dataclass
class Problem:
    P: sp.csc_matrix  # upper triangle
    q: np.ndarray
    A: sp.csc_matrix
    b: np.ndarray
    zero_dim: int = 0
    nonneg_dim: int = 0
    soc_dims: list = field(default_factory=list)
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return self.P.shape[0]

    @property
    def m(self):
        return self.A.shape[0]

    def P_full(self):
        U = self.P.tocsc()
        return (U + sp.triu(U, 1).T).tocsc()


def _rng(config, index=0, timestep=0):
    return np.random.Generator(np.random.PCG64(np.random.SeedSequence([config, index, timestep])))


def _csc(M):
    M = sp.csc_matrix(M)
    M.sum_duplicates()
    M.sort_indices()
    return M


def _upper(Pfull):
    return _csc(sp.triu(sp.csc_matrix(Pfull)))


def _interior(rng, zero_dim, nonneg_dim, soc_dims, zero_kind):
    parts = []
    if zero_dim:
        parts.append(np.zeros(zero_dim) if zero_kind == "zero" else rng.standard_normal(zero_dim))
    if nonneg_dim:
        parts.append(rng.uniform(0.1, 1.0, nonneg_dim))
    for d in soc_dims:
        u = rng.standard_normal(d - 1)
        parts.append(np.concatenate([[np.linalg.norm(u) + rng.uniform(0.1, 1.0)], u]))
    return np.concatenate(parts) if parts else np.zeros(0)


def feasible_qb(rng, Pu, A, zero_dim, nonneg_dim, soc_dims):
    """(q, b, x0, s0, z0) from the seeded strictly feasible primal-dual pair."""
    n = Pu.shape[0]
    x0 = rng.standard_normal(n)
    s0 = _interior(rng, zero_dim, nonneg_dim, soc_dims, "zero")
    z0 = _interior(rng, zero_dim, nonneg_dim, soc_dims, "normal")
    Pf = Pu + sp.triu(Pu, 1).T
    b = A @ x0 + s0
    q = -(Pf @ x0) - A.T @ z0
    return np.asarray(q), np.asarray(b), x0, s0, z0


def _rows_with_k_distinct_cols(rng, m, n, k):
    cols = rng.integers(0, n, size=(m, k))
    while True:
        srt = np.sort(cols, axis=1)
        bad = np.nonzero((np.diff(srt, axis=1) == 0).any(axis=1))[0]
        if bad.size == 0:
            return cols
        cols[bad] = rng.integers(0, n, size=(bad.size, k))


def random_sparse_A(rng, m, n, k=5):
    cols = _rows_with_k_distinct_cols(rng, m, n, k)
    vals = rng.standard_normal((m, k))
    rows = np.repeat(np.arange(m), k)
    return _csc(sp.coo_matrix((vals.ravel(), (rows, cols.ravel())), shape=(m, n)))


# --------------------------------------------------------------------------
# Config 0: n=200, m = 20 zero + 100 nonneg + 20 x Q3 (the CPU-runnable case)
# --------------------------------------------------------------------------
def config0(seed=0, n=200, zero=20, nonneg=100, n_soc=20):
    rng = _rng(0, seed)
    rows_M = n // 2
    pos = rng.choice(rows_M * n, size=int(0.1 * rows_M * n), replace=False)
    M = sp.coo_matrix((rng.standard_normal(pos.size), (pos // n, pos % n)), shape=(rows_M, n))
    Pf = (M.T @ M + 1e-2 * sp.identity(n)).tocsc()
    Pu = _upper(Pf)
    soc = [3] * n_soc
    m = zero + nonneg + 3 * n_soc
    A = random_sparse_A(rng, m, n, 5)
    q, b, *_ = feasible_qb(rng, Pu, A, zero, nonneg, soc)
    return Problem(Pu, q, A, b, zero, nonneg, soc, name=f"c0-n{n}-s{seed}")


# --------------------------------------------------------------------------
# Config 1: n=20000, m = 2000 zero + 20000 nonneg + 2000 Q3 + 10 Q50
# P = diag(U[0.1,1]) + tridiagonal N(0, 0.05); the diagonal is raised to
# max(d_i, |e_{i-1}| + |e_i| + 0.01) (diagonal dominance => P positive
# definite; as written the matrix is indefinite and the QP not convex).
# --------------------------------------------------------------------------
def config1(seed=0, scale=1.0):
    n = int(round(20000 * scale))
    zero, nonneg, n_q3, n_q50 = (int(round(v * scale)) for v in (2000, 20000, 2000, 10))
    n_q50 = max(n_q50, 1)
    rng = _rng(1, seed)
    dg = rng.uniform(0.1, 1.0, n)
    e = rng.normal(0.0, 0.05, n - 1)
    ae = np.abs(e)
    left = np.concatenate([[0.0], ae])
    right = np.concatenate([ae, [0.0]])
    dg = np.maximum(dg, left + right + 0.01)
    Pf = sp.diags([e, dg, e], [-1, 0, 1], shape=(n, n), format="csc")
    Pu = _upper(Pf)
    soc = [3] * n_q3 + [50] * n_q50
    m = zero + nonneg + sum(soc)
    A = random_sparse_A(rng, m, n, 5)
    q, b, *_ = feasible_qb(rng, Pu, A, zero, nonneg, soc)
    return Problem(Pu, q, A, b, zero, nonneg, soc, name=f"c1-x{scale}-s{seed}",
                   meta={"psd_repair": "d_i <- max(d_i, |e_{i-1}|+|e_i|+0.01)"})


# --------------------------------------------------------------------------
# Config 2: one friction-contact QP of the 4096-problem batch.
# c ~ U{4..64} contacts, n = 3c; P = L L' + 1e-3 I with L lower banded
# (diagonal + 6 sub-diagonals, N(0,1)); q ~ N(0,1); per contact one Q3 block
# with rows [-mu a_n; -a_t1; -a_t2] over the contact's 3 variables plus 3
# distinct random other variables; box |x_i| <= 10 as 2n nonnegative rows
# (x_i <= 10 and -x_i <= 10).  Timestep t perturbs the base q and b by
# N(0, 0.01) (not cumulative).
# --------------------------------------------------------------------------
def config2_problem(index, timestep=0, c=None):
    rng = _rng(2, index, 0)
    if c is None:
        c = int(rng.integers(4, 65))
    n = 3 * c
    band = 6
    Ld = np.zeros((n, n))
    for k in range(band + 1):
        Ld += np.diag(rng.standard_normal(n - k), -k)
    Pf = Ld @ Ld.T + 1e-3 * np.eye(n)
    Pu = _upper(Pf)
    q = rng.standard_normal(n)
    mu = rng.uniform(0.2, 1.0, c)
    rows, cols, vals = [], [], []
    nbox = 2 * n
    for i in range(n):
        rows += [i, n + i]
        cols += [i, i]
        vals += [1.0, -1.0]
    for k in range(c):
        own = [3 * k, 3 * k + 1, 3 * k + 2]
        others = rng.choice(n - 3, size=3, replace=False)
        others = [o if o < 3 * k else o + 3 for o in others]
        vs = own + sorted(others)
        a = rng.standard_normal((3, 6))
        a[0] *= mu[k]
        for r in range(3):
            for j in range(6):
                rows.append(nbox + 3 * k + r)
                cols.append(vs[j])
                vals.append(-a[r, j])
    m = nbox + 3 * c
    A = _csc(sp.coo_matrix((vals, (rows, cols)), shape=(m, n)))
    soc = [3] * c
    x0 = rng.standard_normal(n)
    x0 = np.clip(x0, -9.0, 9.0)
    s0 = _interior(rng, 0, 0, soc, "zero")
    b = np.concatenate([np.full(nbox, 10.0), A[nbox:] @ x0 + s0])
    if timestep:
        pr = _rng(2, index, timestep)
        q = q + pr.normal(0.0, 0.01, n)
        b = b + pr.normal(0.0, 0.01, m)
    return Problem(Pu, q, A, b, 0, nbox, soc, name=f"c2-p{index}-t{timestep}",
                   meta={"contacts": c, "box_rows": "2n"})


def config2_perturbation(index, timestep, n, m):
    """The (q, b) perturbation config2_problem(index, timestep) applies to the
    base problem's q and b (N(0, 0.01), seeded per problem and timestep)."""
    pr = _rng(2, index, timestep)
    return pr.normal(0.0, 0.01, n), pr.normal(0.0, 0.01, m)


# --------------------------------------------------------------------------
# Config 3: differentiation batch member: n=64, m = 16 zero + 64 nonneg + 8 Q4,
# P = M'M + 1e-2 I with M 32x64 dense, A dense 112x64, loss weights w ~ N(0,1).
# --------------------------------------------------------------------------
def config3_problem(index):
    rng = _rng(3, index)
    n, zero, nonneg, soc = 64, 16, 64, [4] * 8
    M = rng.standard_normal((32, n))
    Pu = _upper(M.T @ M + 1e-2 * np.eye(n))
    m = zero + nonneg + sum(soc)
    A = _csc(rng.standard_normal((m, n)))
    q, b, *_ = feasible_qb(rng, Pu, A, zero, nonneg, soc)
    w = rng.standard_normal(n)
    return Problem(Pu, q, A, b, zero, nonneg, soc, name=f"c3-p{index}", meta={"w": w})


# --------------------------------------------------------------------------
# Config 4: n=100000, m = 10000 zero + 100000 nonneg + 10000 Q3 + 50 Q100;
# P = diag(U[0.1,1]) + 0.1 * (5-point Laplacian of a 250x400 grid).
# --------------------------------------------------------------------------
def config4(seed=0, scale=1.0):
    gx, gy = 250, 400
    if scale != 1.0:
        f = np.sqrt(scale)
        gx, gy = max(2, int(round(250 * f))), max(2, int(round(400 * f)))
    n = gx * gy
    zero, nonneg, n_q3, n_q100 = (int(round(v * n / 100000)) for v in (10000, 100000, 10000, 50))
    n_q100 = max(n_q100, 1)
    rng = _rng(4, seed)
    Tx = sp.diags([-np.ones(gx - 1), 2 * np.ones(gx), -np.ones(gx - 1)], [-1, 0, 1])
    Ty = sp.diags([-np.ones(gy - 1), 2 * np.ones(gy), -np.ones(gy - 1)], [-1, 0, 1])
    Lap = sp.kron(sp.identity(gy), Tx) + sp.kron(Ty, sp.identity(gx))
    Pf = (0.1 * Lap + sp.diags(rng.uniform(0.1, 1.0, n))).tocsc()
    Pu = _upper(Pf)
    soc = [3] * n_q3 + [100] * n_q100
    m = zero + nonneg + sum(soc)
    A = random_sparse_A(rng, m, n, 5)
    q, b, *_ = feasible_qb(rng, Pu, A, zero, nonneg, soc)
    return Problem(Pu, q, A, b, zero, nonneg, soc, name=f"c4-x{scale}-s{seed}")


# --------------------------------------------------------------------------
# Known-answer problems of the reference's own solver tests.
# --------------------------------------------------------------------------
def _sp(rows, cols, vals, shape):
    return _csc(sp.coo_matrix((vals, (rows, cols)), shape=shape))


# --------------------------------------------------------------------------
# Planted optimum: a strictly convex, nondegenerate conic QP whose optimum
# (x*, s*, z*) is chosen first.  Each cone row or block is given one of the
# strictly complementary patterns
#   nonnegative row:  s* > 0, z* = 0   or   s* = 0, z* > 0
#   SOC block:        s* interior, z* = 0;   s* = 0, z* interior;   or both on
#                     the boundary, s* = a(1, u), z* = c(1, -u), |u| = 1
#                     (s*'z* = 0 and s* + z* interior)
# and q = -(P x* + A'z*), b = A x* + s*.  P is positive definite and the active
# rows of A are random and fewer than n, so the KKT conditions have exactly one
# solution and it is (x*, s*, z*): the optimum is known in closed form and the
# solution map is Lipschitz near it (no degenerate directions that the
# stopping tolerance leaves undetermined).
# --------------------------------------------------------------------------
def planted(seed=0, n=120, zero=10, nonneg=80, soc_dims=(3, 3, 4, 5, 6, 8), p_kind="sparse",
            boundary_pairs=False):
    """p_kind: "sparse" (M'M + I/2, M with n/2 rows of 10% density), "dense"
    (B B'/n + I/2) or "tridiag" (diagonally dominant tridiagonal, for large n).
    boundary_pairs adds SOC blocks with both s* and z* on the boundary; there
    the interior-point iterates approach the optimum only as sqrt(mu)."""
    rng = _rng(100, seed)
    if p_kind == "dense":
        B = rng.standard_normal((n, n)) / np.sqrt(n)
        Pf = sp.csc_matrix(B @ B.T + 0.5 * np.eye(n))
    elif p_kind == "tridiag":
        e = rng.normal(0.0, 0.1, n - 1)
        ae = np.abs(e)
        dg = rng.uniform(0.5, 1.0, n) + np.concatenate([[0.0], ae]) + np.concatenate([ae, [0.0]])
        Pf = sp.diags([e, dg, e], [-1, 0, 1], shape=(n, n), format="csc")
    else:
        rows_M = n // 2
        pos = rng.choice(rows_M * n, size=int(0.1 * rows_M * n), replace=False)
        M = sp.coo_matrix((rng.standard_normal(pos.size), (pos // n, pos % n)), shape=(rows_M, n))
        Pf = (M.T @ M + 0.5 * sp.identity(n)).tocsc()
    Pu = _upper(Pf)
    soc_dims = list(soc_dims)
    m = zero + nonneg + sum(soc_dims)
    A = random_sparse_A(rng, m, n, 5)
    x = rng.standard_normal(n)
    s = np.zeros(m)
    z = np.zeros(m)
    z[:zero] = rng.standard_normal(zero)
    act = rng.random(nonneg) < 0.5
    v = rng.uniform(0.5, 2.0, nonneg)
    s[zero:zero + nonneg] = np.where(act, 0.0, v)
    z[zero:zero + nonneg] = np.where(act, v, 0.0)
    o = zero + nonneg
    for k, d in enumerate(soc_dims):
        u = rng.standard_normal(d - 1)
        u /= np.linalg.norm(u)
        kind = k % 3 if boundary_pairs else k % 2
        if kind == 0:    # s interior, z = 0
            s[o:o + d] = np.concatenate([[rng.uniform(1.5, 2.5)], rng.uniform(0.2, 0.8) * u])
        elif kind == 1:  # s = 0, z interior
            z[o:o + d] = np.concatenate([[rng.uniform(1.5, 2.5)], rng.uniform(0.2, 0.8) * u])
        else:            # both on the boundary, complementary
            s[o:o + d] = rng.uniform(0.5, 2.0) * np.concatenate([[1.0], u])
            z[o:o + d] = rng.uniform(0.5, 2.0) * np.concatenate([[1.0], -u])
        o += d
    b = A @ x + s
    q = -(Pf @ x) - A.T @ z
    return Problem(Pu, np.asarray(q), A, np.asarray(b), zero, nonneg, soc_dims,
                   name=f"planted-n{n}-s{seed}", meta={"x": x, "s": s, "z": z})


def planted_suite():
    """The planted-optimum instances of the tight-tolerance parity tier: small
    sparse and dense P, a 400-variable instance with 20 SOC blocks, and a
    1000-variable tridiagonal-P instance whose KKT factorization has fronts of
    several hundred rows."""
    return [
        planted(0),
        planted(1, p_kind="dense"),
        planted(2, n=400, zero=30, nonneg=300, soc_dims=[3] * 20 + [10, 20]),
        planted(3, n=120, zero=0, nonneg=150, soc_dims=[]),
        planted(4, n=60, zero=5, nonneg=0, soc_dims=[3, 4, 5, 6, 7, 8, 9, 10], p_kind="dense"),
        planted(7, n=1000, zero=100, nonneg=800, soc_dims=[3] * 80 + [50] * 2, p_kind="tridiag"),
    ]


def kat_problems():
    """(name, Problem, expected_status, expected_x or None, expected_obj or None, tol)."""
    out = []
    # test/solver/test_solve_qp.cpp:9-34
    P = _sp([0, 1], [0, 1], [1.0, 1.0], (2, 2))
    A = _sp([0, 0], [0, 1], [1.0, 1.0], (1, 2))
    out.append(("qp_equality", Problem(P, np.zeros(2), A, np.array([1.0]), zero_dim=1),
                1, np.array([0.5, 0.5]), 0.25, 1e-6))
    # test/solver/test_solve_qp.cpp:36-66
    A = _sp([0, 1], [0, 1], [-1.0, -1.0], (2, 2))
    out.append(("qp_box_active", Problem(P, np.array([1.0, -2.0]), A, np.zeros(2), nonneg_dim=2),
                1, np.array([0.0, 2.0]), -2.0, 1e-6))
    # test/solver/test_solve_lp.cpp:9-37
    A = _sp([0, 0, 1, 2], [0, 1, 0, 1], [1.0, 1.0, -1.0, -1.0], (3, 2))
    out.append(("lp_vertex", Problem(_csc(sp.csc_matrix((2, 2))), np.array([-1.0, -2.0]), A,
                                     np.array([1.0, 0.0, 0.0]), nonneg_dim=3),
                1, np.array([0.0, 1.0]), -2.0, 1e-6))
    # test/solver/test_solve_socp.cpp:11-42
    A = _sp([0, 1, 2, 3], [1, 0, 1, 2], [1.0, -1.0, -1.0, -1.0], (4, 3))
    out.append(("socp_min_x0", Problem(_csc(sp.csc_matrix((3, 3))), np.array([1.0, 0.0, 0.0]), A,
                                       np.array([1.0, 0, 0, 0]), zero_dim=1, soc_dims=[3]),
                1, np.array([1.0, 1.0, 0.0]), 1.0, 1e-5))
    # test/solver/test_solve_friction_qp_soc.cpp:17-51 (projection of (1,3,4) onto Q3)
    f = np.array([1.0, 3.0, 4.0])
    a = np.linalg.norm(f[1:])
    proj = np.concatenate([[0.5 * (f[0] + a)], (0.5 * (f[0] + a) / a) * f[1:]])
    P3 = _sp([0, 1, 2], [0, 1, 2], [1.0] * 3, (3, 3))
    A3 = _sp([0, 1, 2], [0, 1, 2], [-1.0] * 3, (3, 3))
    out.append(("soc_projection", Problem(P3, -f, A3, np.zeros(3), soc_dims=[3]),
                1, proj, None, 1e-5))
    # README.md:74-121: min 1/2||x-c||^2, c=(1,1,0.3), x0+x1=1, x2>=0, (x2,x0,x1) in Q3
    A = _sp([0, 0, 1, 2, 3, 4], [0, 1, 2, 2, 0, 1], [1.0, 1.0, -1.0, -1.0, -1.0, -1.0], (5, 3))
    out.append(("readme_three_cones", Problem(P3, np.array([-1.0, -1.0, -0.3]), A,
                                              np.array([1.0, 0, 0, 0, 0]), zero_dim=1,
                                              nonneg_dim=1, soc_dims=[3]),
                1, np.array([0.5, 0.5, np.sqrt(0.5)]), None, 1e-5))
    # test/solver/test_infeasible.cpp:9-31
    A = _sp([0, 1], [0, 0], [-1.0, 1.0], (2, 1))
    out.append(("primal_infeasible", Problem(_csc(sp.csc_matrix((1, 1))), np.zeros(1), A,
                                             np.array([-1.0, 0.0]), nonneg_dim=2),
                2, None, None, None))
    # test/solver/test_infeasible.cpp:33-50
    A = _sp([0], [0], [-1.0], (1, 1))
    out.append(("dual_infeasible", Problem(_csc(sp.csc_matrix((1, 1))), np.array([-1.0]), A,
                                           np.zeros(1), nonneg_dim=1),
                3, None, None, None))
    return out


def friction_demo_problem(f_des, mu=0.5):
    """examples/friction_contact_demo.cpp:19-71: min 1/2||f||^2 - f_des'f, -diag(mu,1,1) f in Q3."""
    P = _sp([0, 1, 2], [0, 1, 2], [1.0] * 3, (3, 3))
    A = _sp([0, 1, 2], [0, 1, 2], [-mu, -1.0, -1.0], (3, 3))
    return Problem(P, -np.asarray(f_des, float), A, np.zeros(3), soc_dims=[3])


def box_qp(c):
    """test/solver/test_update_reuse.cpp:17-30: min 1/2||x-c||^2 s.t. x >= 0."""
    n = len(c)
    P = _sp(list(range(n)), list(range(n)), [1.0] * n, (n, n))
    A = _sp(list(range(n)), list(range(n)), [-1.0] * n, (n, n))
    return Problem(P, -np.asarray(c, float), A, np.zeros(n), nonneg_dim=n)


def kkt_certificate(prob, x, s, z):
    """Optimality certificate (test/utils/test_helpers.h:32-48): residual norms and margins."""
    Pf = prob.P_full()
    st = Pf @ x + prob.A.T @ z + prob.q
    fe = prob.A @ x + s - prob.b
    out = dict(stationarity=float(np.linalg.norm(st)), feasibility=float(np.linalg.norm(fe)),
               complementarity=float(abs(s @ z)))
    off = prob.zero_dim
    mins, minz = np.inf, np.inf
    if prob.nonneg_dim:
        mins = min(mins, s[off:off + prob.nonneg_dim].min())
        minz = min(minz, z[off:off + prob.nonneg_dim].min())
        off += prob.nonneg_dim
    for d in prob.soc_dims:
        mins = min(mins, s[off] - np.linalg.norm(s[off + 1:off + d]))
        minz = min(minz, z[off] - np.linalg.norm(z[off + 1:off + d]))
        off += d
    out["margin_s"] = float(mins if np.isfinite(mins) else 0.0)
    out["margin_z"] = float(minz if np.isfinite(minz) else 0.0)
    return out


def kkt_lower_pattern(prob):
    """Lower-triangle pattern of the reference's KKT matrix (kkt_system.cpp:54-107)."""
    n, m = prob.n, prob.m
    rows, cols = [np.arange(n)], [np.arange(n)]
    Pc = prob.P.tocoo()
    sel = Pc.row < Pc.col
    rows.append(Pc.col[sel]); cols.append(Pc.row[sel])
    Ac = prob.A.tocoo()
    rows.append(n + Ac.row); cols.append(Ac.col)
    rows.append(n + np.arange(prob.zero_dim + prob.nonneg_dim))
    cols.append(n + np.arange(prob.zero_dim + prob.nonneg_dim))
    off = prob.zero_dim + prob.nonneg_dim
    for d in prob.soc_dims:
        a, b = np.tril_indices(d)
        rows.append(n + off + a); cols.append(n + off + b)
        off += d
    r = np.concatenate(rows); c = np.concatenate(cols)
    K = sp.csc_matrix((np.ones(r.size), (r, c)), shape=(n + m, n + m))
    K.sum_duplicates()
    K.sort_indices()
    K.data[:] = 1.0
    return K
