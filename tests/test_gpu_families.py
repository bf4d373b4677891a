"""The benchmark-family suite on the GPU (conicxx_amd.families; reference
benchmarks/problem_generators.cpp:96-405, run_benchmarks.cpp:132-188): every
instance must end Solved like the oracle, within the recorded GPU iteration
count plus one (the iteration regression gate, tests/golden/
family_iterations.json) and within three iterations of the oracle's count, at
the oracle's objective, with the KKT certificate; on the strictly convex
families (RandomQP, FrictionChain, FrictionChainXL: P positive definite, a
unique optimum) the solution x must agree with the oracle's."""
import json
import os

import numpy as np
import pytest

from conicxx_amd import ConeSpec, Settings, Solver, Status
from conicxx_amd import families as F
from conicxx_amd import problems as PR

pytestmark = pytest.mark.gpu

FIXTURE = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "family_iterations.json")))
STRICTLY_CONVEX = ("RandomQP", "FrictionChain", "FrictionChainXL")


@pytest.mark.parametrize("key", sorted(FIXTURE))
def test_family_instance_matches_oracle(key, oracle):
    fam, size, trial = key.split("/")
    p = F.make(fam, int(size), int(trial))
    s = Solver()
    assert s.setup(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims), Settings())
    sol = s.solve()
    r = oracle.solve(p.P, p.q, p.A, p.b, p.zero_dim, p.nonneg_dim, p.soc_dims)
    rec = FIXTURE[key]
    print(key, sol.status.name, sol.info.iterations, "oracle", r["iterations"], "recorded", rec["gpu_iterations"])
    assert int(sol.status) == r["status"] == int(Status.Solved)
    assert sol.info.iterations <= rec["gpu_iterations"] + 1
    assert abs(sol.info.iterations - r["iterations"]) <= 3
    # objectives at the 1e-8 relative-gap tolerance
    assert abs(sol.objective - r["objective"]) <= 1e-7 * (1.0 + abs(r["objective"])), (sol.objective, r["objective"])
    c = PR.kkt_certificate(p, sol.x, sol.s, sol.z)
    sc = 1.0 + np.abs(p.q).max() + np.abs(p.b).max()
    assert c["stationarity"] < 1e-6 * sc and c["feasibility"] < 1e-6 * sc, c
    assert c["margin_s"] > -1e-8 * sc and c["margin_z"] > -1e-8 * sc, c
    if fam in STRICTLY_CONVEX:
        err = np.linalg.norm(sol.x - r["x"]) / max(1.0, np.linalg.norm(r["x"]))
        assert err < 1e-6, err
