"""CPU tests: the C-ABI library loads and exports every declared symbol, and the
host symbolic analysis (ordering, elimination tree, column counts, supernode
plan) is exact.  No compute call touches a GPU here."""
import ctypes
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

from conicxx_amd import _lib, analyze_pattern
from conicxx_amd import problems as PR

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "conicxx_c.h")).read()
    declared = set(re.findall(r"\b(cx_[a-z_0-9]+)\s*\(", hdr))
    assert declared, "no declarations found"
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(L, name), f"{name} declared in include/conicxx_c.h but not exported"
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert declared == bound


def test_default_settings_match_reference():
    st = _lib.CxSettings()
    _lib.lib().cx_default_settings(ctypes.byref(st))
    # include/conicxx/settings.h:23-61
    assert (st.tol_feas, st.tol_gap_abs, st.tol_gap_rel, st.tol_infeas) == (1e-8, 1e-8, 1e-8, 1e-7)
    assert st.max_iter == 200 and st.refine_max_iter == 3 and st.refine_tol == 1e-12
    assert (st.static_reg_P, st.static_reg_A) == (1e-8, 1e-8)
    assert (st.dynamic_reg_eps, st.dynamic_reg_delta) == (1e-14, 1e-7)
    assert st.max_step_fraction == 0.99 and st.equilibrate == 1 and st.equilibrate_max_iter == 10
    assert (st.equilibrate_min_scale, st.equilibrate_max_scale) == (1e-4, 1e4)
    assert st.warm_start == 1 and st.linear_solver == 1


def test_status_strings():
    names = ["Unsolved", "Solved", "PrimalInfeasible", "DualInfeasible", "MaxIterations",
             "NumericalError"]
    for i, nm in enumerate(names):
        assert _lib.lib().cx_status_string(i).decode() == nm


def _brute_colcounts(K_lower, perm):
    """Column counts of L for P K P' by explicit symbolic elimination."""
    N = K_lower.shape[0]
    iperm = np.empty(N, dtype=int)
    iperm[perm] = np.arange(N)
    K = sp.coo_matrix(K_lower)
    adj = [set() for _ in range(N)]
    for r, c in zip(K.row, K.col):
        if r == c:
            continue
        a, b = iperm[r], iperm[c]
        adj[min(a, b)].add(max(a, b))
    cc = np.zeros(N, dtype=int)
    for j in range(N):
        s = adj[j]
        cc[j] = len(s) + 1
        if s:
            p = min(s)
            adj[p] |= (s - {p})
    return cc


# c1x0.1 has thin fronts (a few pivot columns over an update matrix of more
# than 1024 rows) that the second elimination order absorbs into their parents:
# the counts of that final order are checked too.
@pytest.mark.parametrize("maker", [lambda: PR.config0(), lambda: PR.config2_problem(3),
                                   lambda: PR.config3_problem(1), lambda: PR.config1(scale=0.05),
                                   lambda: PR.config1(scale=0.1), lambda: PR.config4(scale=0.01)],
                         ids=["c0", "c2", "c3", "c1x0.05", "c1x0.1", "c4x0.01"])
def test_symbolic_column_counts_exact(maker):
    p = maker()
    K = PR.kkt_lower_pattern(p)
    perm, cc, st = analyze_pattern(K)
    N = K.shape[0]
    assert sorted(perm.tolist()) == list(range(N))
    ref = _brute_colcounts(K, perm)
    np.testing.assert_array_equal(cc, ref)
    assert st["nnz_L_exact"] == ref.sum()
    assert st["nnz_L"] >= st["nnz_L_exact"]  # relaxed supernodes store extra zeros
    assert st["nnz_L"] <= 1.5 * st["nnz_L_exact"] + 1e4


def test_ordering_quality_comparable_to_exact_min_degree(oracle):
    """AMD fill within 25% of the oracle's exact minimum-degree fill."""
    p = PR.config1(scale=0.05)
    K = PR.kkt_lower_pattern(p)
    _, _, st = analyze_pattern(K)
    s = oracle.OracleSolver()
    assert s.setup(p.P, p.q, p.A, p.b, p.zero_dim, p.nonneg_dim, p.soc_dims)
    nnz_oracle = s.kkt_stats()["nnz_L"]
    assert st["nnz_L_exact"] < 1.25 * nnz_oracle


@pytest.mark.parametrize("maker", [lambda: PR.config0(seed=2), lambda: PR.config2_problem(7),
                                   lambda: PR.config3_problem(1), lambda: PR.config1(scale=0.02)],
                         ids=["config0", "contact", "dense", "config1"])
@pytest.mark.parametrize("st", [{}, {"equilibrate_max_iter": 25}, {"equilibrate_max_scale": 2.0}],
                         ids=["default", "iters25", "maxscale2"])
def test_library_equilibration_matches_oracle(maker, st, oracle):
    """The library's host Ruiz equilibration (what setup() scales by) against
    the oracle's restatement of equilibration.cpp:8-99, no GPU needed."""
    from conicxx_amd import ConeSpec, Settings, equilibrate
    p = maker()
    d, e, c = equilibrate(p.P, p.A, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims), Settings(**st))
    do, eo, co = oracle.equilibrate(p.P, p.A, p.zero_dim, p.nonneg_dim, p.soc_dims, **st)
    np.testing.assert_allclose(d, do, rtol=1e-14, atol=0)
    np.testing.assert_allclose(e, eo, rtol=1e-14, atol=0)
    assert abs(c - co) <= 1e-14 * abs(co)
