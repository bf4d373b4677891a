"""The benchmark-family generators (conicxx_amd.families) on the CPU: the
structure of each family (benchmarks/problem_generators.cpp:96-405) and the
oracle solving every instance of the suite (run_benchmarks.cpp:132-188) to
Solved within the recorded iteration counts (tests/golden/family_iterations.json,
written by tests/golden/make_family_iterations.py)."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

from conicxx_amd import families as F
from conicxx_amd import problems as PR

FIXTURE = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "family_iterations.json")))


def _check_common(p):
    assert p.zero_dim + p.nonneg_dim + sum(p.soc_dims) == p.m
    assert p.P.shape == (p.n, p.n) and p.A.shape[1] == p.n and len(p.q) == p.n and len(p.b) == p.m
    assert sp.tril(p.P, -1).nnz == 0, "P must hold the upper triangle only"
    ev = np.linalg.eigvalsh(p.P_full().toarray()) if p.n <= 400 else None
    if ev is not None:
        assert ev.min() > -1e-9, "P must be positive semidefinite"


def test_suite_composition():
    """Sizes and trial counts of the reference runner: 62 instances."""
    assert sum(len(s) * t for s, t in F.SUITE.values()) == 62
    keys = [p.meta["key"] for p in F.instances()]
    assert len(keys) == 62 and set(keys) == set(FIXTURE)
    assert len(list(F.instances(max_xl=66))) == 60


def test_random_qp_lp_socp_structure():
    p = F.make("RandomQP", 10, 0)
    _check_common(p)
    assert p.nonneg_dim == 20 and np.allclose(p.A.toarray(), np.vstack([-np.eye(10), np.eye(10)]))
    assert np.linalg.eigvalsh(p.P_full().toarray()).min() >= 1e-2 - 1e-12
    lp = F.make("RandomLP", 10, 0)
    _check_common(lp)
    assert lp.P.nnz == 0 and lp.nonneg_dim == 20
    s = F.make("RandomSOCP", 5, 0)
    _check_common(s)
    assert s.soc_dims == [4] * 5 and s.n == s.m == 20 and s.A.nnz == 400 and s.P.nnz == 0


def test_portfolio_structure():
    p = F.make("Portfolio", 20, 0)
    _check_common(p)
    k = 3
    assert p.n == 24 and p.zero_dim == 1 and p.nonneg_dim == 20 and p.soc_dims == [21, k + 1, 3, 3]
    A = p.A.toarray()
    assert np.array_equal(A[0, :20], np.ones(20)) and p.b[0] == 1.0          # budget 1'x = 1
    assert np.array_equal(A[1:21, :20], -np.eye(20))                         # x >= 0
    assert p.q[21] == 5.0 and p.q[20] == 5.0 and np.all(p.q[:20] <= 0)     # -mu, gamma on t and s
    assert np.array_equal(p.b[-6:], [1.0, 1.0, 0.0, 1.0, 1.0, 0.0])        # rotated-cone rows


def test_group_lasso_and_friction_structure():
    g = F.make("GroupLasso", 10, 0)
    _check_common(g)
    assert g.soc_dims == [6] * 10 and g.n == 60 and np.all(g.b == 0) and np.all(g.q[50:] == 0.1)
    f = F.make("FrictionChain", 4, 0)
    _check_common(f)
    assert f.zero_dim == 1 and f.soc_dims == [3] * 4 and (f.P != sp.identity(12)).nnz == 0
    A = f.A.toarray()
    assert np.array_equal(A[0], np.tile([1.0, 0.0, 0.0], 4))
    assert A[1, 0] == -0.7 and A[2, 1] == -1.0 and A[3, 2] == -1.0
    assert np.isclose(f.b[0], -1.2 * f.q[0::3].sum())
    x = F.make("FrictionChainXL", 8, 0)
    _check_common(x)
    assert x.n == 600 and x.zero_dim == 600 and x.soc_dims == [3] * 8
    assert np.all(np.diff(x.A[:600].tocsr().indptr) == 3)                  # banded chain rows


def test_deterministic():
    a, b = F.make("GroupLasso", 10, 1), F.make("GroupLasso", 10, 1)
    assert np.array_equal(a.q, b.q) and (a.P != b.P).nnz == 0
    assert not np.array_equal(F.make("GroupLasso", 10, 2).q, a.q)


@pytest.mark.parametrize("key", sorted(FIXTURE))
def test_oracle_solves_every_instance(key, oracle):
    """The oracle reaches Solved on every instance, within two iterations of the
    recorded count (the iteration regression gate; rounding-level differences
    of the host's libm can move a count by one or two), at the recorded
    objective, with the KKT certificate."""
    fam, size, trial = key.split("/")
    p = F.make(fam, int(size), int(trial))
    r = oracle.solve(p.P, p.q, p.A, p.b, p.zero_dim, p.nonneg_dim, p.soc_dims)
    rec = FIXTURE[key]
    assert r["status"] == rec["oracle_status"] == 1
    assert r["iterations"] <= rec["oracle_iterations"] + 2, (r["iterations"], rec)
    assert abs(r["objective"] - rec["oracle_objective"]) <= 1e-7 * (1.0 + abs(rec["oracle_objective"]))
    c = PR.kkt_certificate(p, r["x"], r["s"], r["z"])
    sc = 1.0 + np.abs(p.q).max() + np.abs(p.b).max()
    assert c["stationarity"] < 1e-6 * sc and c["feasibility"] < 1e-6 * sc, c
