"""Precision form on the GPU: FP32 numeric factorization (FP32 fronts, FP32
MFMA trailing updates) as the preconditioner of FP64 iterative refinement,
with an FP64 fallback when refinement misses refine_tol.  Same parity bar as
the FP64 tests against the CPU oracle (which is FP64 throughout)."""
import numpy as np
import pytest

from conicxx_amd import ConeSpec, FactorPrecision, Settings, Solver, Status
from conicxx_amd import problems as PR
from test_gpu_solver import assert_parity, oracle_runs

pytestmark = pytest.mark.gpu

FP32 = Settings(factor_precision=FactorPrecision.FP32Refined)


def solve(p, settings):
    s = Solver()
    assert s.setup(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims), settings)
    return s, s.solve()


@pytest.mark.parametrize("seed", [0, 1])
def test_fp32_config0_parity(seed, oracle):
    p = PR.config0(seed=seed)
    s, sol = solve(p, FP32)
    pr = s.profile()
    print("profile", {k: pr[k] for k in ("factorizations", "fp32_factorizations", "fp64_fallbacks", "kkt_solves")})
    assert pr["fp32_factorizations"] >= 1
    assert_parity(sol, oracle_runs(oracle, p), p)


def test_fp32_large_sparse_parity(oracle):
    """config 1 scaled down: big fronts on the FP32 MFMA path."""
    p = PR.config1(scale=0.05)
    s, sol = solve(p, FP32)
    pr = s.profile()
    print("profile", {k: pr[k] for k in ("factorizations", "fp32_factorizations", "fp64_fallbacks", "kkt_solves")})
    assert s.kkt_stats()["max_front"] > 256
    assert pr["fp32_factorizations"] >= 1
    assert_parity(sol, oracle_runs(oracle, p), p)


def test_fp32_laplacian_family_parity(oracle):
    """The precision workload's family (grid Laplacian P) at a size the oracle solves in seconds."""
    p = PR.config4(scale=0.01)
    s, sol = solve(p, FP32)
    assert_parity(sol, oracle_runs(oracle, p, npert=2), p)


def test_fp32_and_fp64_agree(oracle):
    """Both GPU precisions inside the oracle's own rounding envelope (two runs
    of the method that differ at the rounding level may end anywhere in the
    1e-8 tolerance ball, ~1e-5 apart in x on this family)."""
    p = PR.config1(scale=0.1)
    _, a = solve(p, Settings())
    _, b = solve(p, FP32)
    assert a.status == b.status == Status.Solved
    err = np.abs(a.x - b.x).max() / max(1.0, np.abs(a.x).max())
    print("FP32+IR vs FP64 x error", err)
    runs = oracle_runs(oracle, p)
    assert_parity(a, runs, p)
    assert_parity(b, runs, p)
