"""Writes family_iterations.json: for every instance of the benchmark-family
suite (conicxx_amd.families), the oracle's status, iteration count and
objective, and, when run with --gpu on a machine with the GPU build, the GPU
solver's status and iteration count (the recorded counts are the iteration
regression gate of tests/test_gpu_families.py).

    python tests/golden/make_family_iterations.py [--gpu] [--out PATH]
"""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpu", action="store_true")
    ap.add_argument("--out", default=os.path.join(HERE, "family_iterations.json"))
    args = ap.parse_args()
    import oracle_py
    from conicxx_amd import families as F
    old = {}
    if os.path.exists(args.out):
        old = json.load(open(args.out))
    rec = {}
    for p in F.instances():
        k = p.meta["key"]
        r = oracle_py.solve(p.P, p.q, p.A, p.b, p.zero_dim, p.nonneg_dim, p.soc_dims)
        rec[k] = {"oracle_status": int(r["status"]), "oracle_iterations": int(r["iterations"]),
                  "oracle_objective": float(r["objective"])}
        if k in old and "gpu_iterations" in old[k]:
            rec[k]["gpu_status"] = old[k]["gpu_status"]
            rec[k]["gpu_iterations"] = old[k]["gpu_iterations"]
        if args.gpu:
            from conicxx_amd import ConeSpec, Settings, Solver
            s = Solver()
            assert s.setup(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims), Settings())
            sol = s.solve()
            rec[k]["gpu_status"] = int(sol.status)
            rec[k]["gpu_iterations"] = int(sol.info.iterations)
        print(k, rec[k], flush=True)
    with open(args.out, "w") as f:
        json.dump(rec, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
