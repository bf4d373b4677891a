"""The reference's benchmark families as own generators, and their runner.

Seven deterministic families of conic problems, each (except Portfolio and
GroupLasso, whose data fix the optimum by construction) built from an explicit
strictly feasible primal-dual pair so every instance has a finite optimum
(benchmarks/problem_generators.cpp:58-75 makeFeasiblePrimalDualData):

  RandomQP      P = M M' + 1e-2 I (M n x n/2, N(0, 1/rank)), box rows A = [-I; I]
                (problem_generators.cpp:96-125)
  RandomLP      the same rows with P = 0 (problem_generators.cpp:127-133)
  RandomSOCP    dense A (N(0, 1/n)), `blocks` SOC blocks of dim 4, P = 0
                (problem_generators.cpp:135-152)
  Portfolio     risk-adjusted return with a factor model: x >= 0, 1'x = 1,
                SOC (u, D^1/2 x), SOC (v, F'x), and the two rotated-cone rows
                (1+t, 1-t, 2u), (1+s, 1-s, 2v) (problem_generators.cpp:154-248)
  GroupLasso    least squares + lambda sum_g t_g with SOC (t_g, x_g) per group
                (problem_generators.cpp:250-300)
  FrictionChain P = I, q = -f_des, one total-load equality, a friction cone per
                contact (problem_generators.cpp:302-346)
  FrictionChainXL 75 banded equality rows per contact over all variables plus
                a friction cone per contact (problem_generators.cpp:348-405)

Sizes and trials are those of the reference runner (run_benchmarks.cpp:
132-188).  The random draws use numpy's PCG64 seeded by (family, size, trial):
the instances have the reference's structure and distributions, not its bit
patterns (its seeds come from std::hash of the family name).

`python -m conicxx_amd.families` solves every instance on the GPU and prints
the reference's table (family, size, trial, status, iterations, time,
residuals), exiting nonzero when an instance is not Solved.
"""
import sys
import time
import zlib
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from .problems import Problem

# family -> (sizes, trials), run_benchmarks.cpp:137-182
SUITE = {
    "RandomQP": ((10, 50, 200), 3),
    "RandomLP": ((10, 50, 200), 3),
    "RandomSOCP": ((5, 20, 50), 3),
    "Portfolio": ((20, 60, 120), 3),
    "GroupLasso": ((10, 30, 60), 3),
    "FrictionChain": ((4, 16, 64), 3),
    "FrictionChainXL": ((8, 33, 66, 264), 2),
}


def family_rng(family, size, trial):
    """Deterministic generator per (family, size, trial) (run_benchmarks.cpp:
    117, problem_generators.cpp:90-94 makeRng)."""
    return np.random.default_rng([zlib.crc32(family.encode()), int(size), int(trial)])


def _upper(M):
    return sp.triu(sp.csc_matrix(M), format="csc")


def _interior(rng, kind, dim):
    """A strictly interior point of one block: 0 on the zero cone, U[0.5, 1.5]
    on the nonnegative orthant, (2, tail of norm 0.6) on a second-order cone
    (problem_generators.cpp:30-53)."""
    if kind == "zero":
        return np.zeros(dim)
    if kind == "nonneg":
        return rng.uniform(0.5, 1.5, dim)
    v = np.empty(dim)
    v[0] = 2.0
    if dim > 1:
        t = rng.uniform(-1.0, 1.0, dim - 1)
        nt = np.linalg.norm(t)
        v[1:] = t * (0.6 / nt) if nt > 1e-12 else t
    return v


def _feasible_qb(P, A, zero_dim, nonneg_dim, soc_dims, rng):
    """q and b from a strictly feasible pair (x0, s0, z0): A x0 + s0 = b and
    P x0 + A' z0 + q = 0 (problem_generators.cpp:58-75)."""
    n = P.shape[0]
    x0 = rng.uniform(-1.0, 1.0, n)
    blocks = [("zero", zero_dim)] * (zero_dim > 0) + [("nonneg", nonneg_dim)] * (nonneg_dim > 0) + \
             [("soc", d) for d in soc_dims]
    s0 = np.zeros(A.shape[0])
    z0 = np.zeros(A.shape[0])
    off = 0
    for kind, d in blocks:
        s0[off:off + d] = _interior(rng, kind, d)
        z0[off:off + d] = _interior(rng, kind, d)
        off += d
    Pf = P + sp.triu(P, 1).T
    return -(Pf @ x0) - A.T @ z0, A @ x0 + s0


def random_qp(n, rng, lp=False):
    rank = max(1, n // 2)
    M = rng.normal(0.0, 1.0 / np.sqrt(rank), (n, rank))
    P = _upper(M @ M.T + 1e-2 * np.eye(n))
    A = sp.vstack([-sp.identity(n), sp.identity(n)], format="csc")
    if lp:
        # RandomLP draws the QP instance first and then replaces P by zero and
        # the feasible pair (problem_generators.cpp:127-133)
        _feasible_qb(P, A, 0, 2 * n, (), rng)
        P = sp.csc_matrix((n, n))
    q, b = _feasible_qb(P, A, 0, 2 * n, (), rng)
    return Problem(P, q, A, b, 0, 2 * n, [], meta={"family": "RandomLP" if lp else "RandomQP", "size": f"n={n}"})


def random_socp(blocks, rng, dim=4):
    n = m = blocks * dim
    A = sp.csc_matrix(rng.normal(0.0, 1.0 / np.sqrt(n), (m, n)))
    P = sp.csc_matrix((n, n))
    soc = [dim] * blocks
    q, b = _feasible_qb(P, A, 0, 0, soc, rng)
    return Problem(P, q, A, b, 0, 0, soc, meta={"family": "RandomSOCP", "size": f"blocks={blocks},dim={dim}"})


def portfolio(assets, rng, factors=None, gamma=5.0):
    n = assets
    k = factors if factors is not None else max(3, assets // 8)
    mu = rng.uniform(0.0, 0.15, n)
    F = rng.normal(0.0, 1.0 / np.sqrt(k), (n, k))
    dsq = rng.uniform(0.05, 0.2, n)
    nx = n + 4
    it, is_, iu, iv = n, n + 1, n + 2, n + 3
    q = np.zeros(nx)
    q[:n] = -mu
    q[it] = gamma
    q[is_] = gamma
    rows, cols, vals, b = [], [], [], []

    def row(entries, rhs):
        r = len(b)
        for c, v in entries:
            rows.append(r), cols.append(c), vals.append(v)
        b.append(rhs)

    row([(i, 1.0) for i in range(n)], 1.0)                      # 1'x = 1
    for i in range(n):                                          # x >= 0
        row([(i, -1.0)], 0.0)
    row([(iu, -1.0)], 0.0)                                      # (u, D^1/2 x) in Q
    for i in range(n):
        row([(i, -dsq[i])], 0.0)
    row([(iv, -1.0)], 0.0)                                      # (v, F'x) in Q
    for j in range(k):
        row([(i, -F[i, j]) for i in range(n) if F[i, j] != 0.0], 0.0)
    row([(it, -1.0)], 1.0), row([(it, 1.0)], 1.0), row([(iu, -2.0)], 0.0)    # (1+t, 1-t, 2u) in Q
    row([(is_, -1.0)], 1.0), row([(is_, 1.0)], 1.0), row([(iv, -2.0)], 0.0)  # (1+s, 1-s, 2v) in Q
    A = sp.csc_matrix((vals, (rows, cols)), shape=(len(b), nx))
    return Problem(sp.csc_matrix((nx, nx)), q, A, np.array(b), 1, n, [n + 1, k + 1, 3, 3],
                   meta={"family": "Portfolio", "size": f"assets={assets},factors={k}"})


def group_lasso(groups, rng, dim=5, measurements=None, lam=0.1):
    n = groups * dim
    nx = n + groups
    mcount = measurements if measurements is not None else 8 * groups
    x_true = np.zeros(n)
    for _ in range(max(1, groups // 4)):
        g = int(rng.integers(0, groups))
        x_true[g * dim:(g + 1) * dim] = rng.uniform(-2.0, 2.0, dim)
    Als = rng.normal(0.0, 1.0 / np.sqrt(n), (mcount, n))
    bls = Als @ x_true + rng.uniform(-0.05, 0.05, mcount)
    Pd = np.zeros((nx, nx))
    Pd[:n, :n] = 2.0 * Als.T @ Als
    q = np.zeros(nx)
    q[:n] = -2.0 * Als.T @ bls
    q[n:] = lam
    rows, cols = [], []
    r = 0
    for g in range(groups):
        rows.append(r), cols.append(n + g)
        r += 1
        for j in range(dim):
            rows.append(r), cols.append(g * dim + j)
            r += 1
    A = sp.csc_matrix((-np.ones(len(rows)), (rows, cols)), shape=(r, nx))
    return Problem(_upper(Pd), q, A, np.zeros(r), 0, 0, [dim + 1] * groups,
                   meta={"family": "GroupLasso", "size": f"groups={groups},dim={dim}"})


def _friction_rows(rows, cols, vals, r0, contacts, mu):
    r = r0
    for c in range(contacts):
        for j, v in ((3 * c, -mu), (3 * c + 1, -1.0), (3 * c + 2, -1.0)):
            rows.append(r), cols.append(j), vals.append(v)
            r += 1
    return r


def friction_chain(contacts, rng, mu=0.7):
    nx = 3 * contacts
    f = np.empty(nx)
    for c in range(contacts):
        f[3 * c] = rng.uniform(5.0, 15.0)
        f[3 * c + 1:3 * c + 3] = rng.uniform(-6.0, 6.0, 2)   # may leave the friction cone
    rows, cols, vals = [], [], []
    for c in range(contacts):                               # sum of normal forces = 1.2 x load
        rows.append(0), cols.append(3 * c), vals.append(1.0)
    r = _friction_rows(rows, cols, vals, 1, contacts, mu)
    A = sp.csc_matrix((vals, (rows, cols)), shape=(r, nx))
    b = np.zeros(r)
    b[0] = 1.2 * f[0::3].sum()
    return Problem(sp.identity(nx, format="csc"), -f, A, b, 1, 0, [3] * contacts,
                   meta={"family": "FrictionChain", "size": f"contacts={contacts}"})


def friction_chain_xl(contacts, rng, mu=0.7, eq_per_contact=75):
    n = eq_per_contact * contacts
    rows, cols, vals = [], [], []
    for i in range(n):                                      # banded equality chain over all variables
        for j, s in ((i, 1.0), ((i + 1) % n, 0.5), ((i + 2) % n, 0.25)):
            rows.append(i), cols.append(j), vals.append(s * rng.uniform(0.5, 1.5))
    r = _friction_rows(rows, cols, vals, n, contacts, mu)
    A = sp.csc_matrix((vals, (rows, cols)), shape=(r, n))   # duplicates (tiny n) are summed
    P = sp.identity(n, format="csc")
    soc = [3] * contacts
    q, b = _feasible_qb(P, A, n, 0, soc, rng)
    return Problem(P, q, A, b, n, 0, soc, meta={"family": "FrictionChainXL", "size": f"contacts={contacts}"})


MAKERS = {
    "RandomQP": lambda s, rng: random_qp(s, rng),
    "RandomLP": lambda s, rng: random_qp(s, rng, lp=True),
    "RandomSOCP": lambda s, rng: random_socp(s, rng),
    "Portfolio": lambda s, rng: portfolio(s, rng),
    "GroupLasso": lambda s, rng: group_lasso(s, rng),
    "FrictionChain": lambda s, rng: friction_chain(s, rng),
    "FrictionChainXL": lambda s, rng: friction_chain_xl(s, rng),
}


def make(family, size, trial):
    p = MAKERS[family](size, family_rng(family, size, trial))
    p.meta.update(trial=trial, key=f"{family}/{size}/{trial}")
    return p


def instances(families=None, max_xl=None):
    """Every (family, size, trial) instance of the suite, in the runner's order."""
    for fam, (sizes, trials) in SUITE.items():
        if families and fam not in families:
            continue
        for size in sizes:
            if fam == "FrictionChainXL" and max_xl is not None and size > max_xl:
                continue
            for t in range(trials):
                yield make(fam, size, t)


@dataclass
class Row:
    key: str
    status: str
    iterations: int
    ms: float
    primal: float
    dual: float
    gap: float
    extra: dict = field(default_factory=dict)


def run(families=None, settings=None, out=sys.stdout):
    """Solve every instance on the GPU and print the reference's table
    (run_benchmarks.cpp:40-107); returns the rows and whether all Solved."""
    from . import ConeSpec, Settings, Solver, Status
    st = settings or Settings()
    rows = []
    print(f"{'Family':<16} {'Size':<24} {'Trial':>5} {'Status':<15} {'Iters':>5} {'Time(ms)':>9} "
          f"{'Primal':>9} {'Dual':>9} {'Gap':>9}", file=out)
    for p in instances(families):
        s = Solver()
        ok = s.setup(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims), st)
        t0 = time.perf_counter()
        sol = s.solve() if ok else None
        ms = 1e3 * (time.perf_counter() - t0)
        r = Row(p.meta["key"], sol.status.name if sol else "SetupFailed", sol.info.iterations if sol else 0, ms,
                sol.info.primal_residual if sol else np.nan, sol.info.dual_residual if sol else np.nan,
                sol.info.duality_gap if sol else np.nan)
        rows.append(r)
        print(f"{p.meta['family']:<16} {p.meta['size']:<24} {p.meta['trial']:>5} {r.status:<15} "
              f"{r.iterations:>5} {ms:>9.2f} {r.primal:>9.2e} {r.dual:>9.2e} {r.gap:>9.2e}", file=out, flush=True)
    all_ok = all(r.status == Status.Solved.name for r in rows)
    print("ALL BENCHMARKS CONVERGED (Status::Solved)." if all_ok else
          "REGRESSION: at least one benchmark instance did not converge.", file=out)
    return rows, all_ok


if __name__ == "__main__":
    _, ok = run(sys.argv[1:] or None)
    sys.exit(0 if ok else 1)
