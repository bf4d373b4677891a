"""Benchmark: QPs solved per second on one MI355X, plain FP64 method.

Workload (default): one large sparse conic QP, n = 20000 variables, m = 28500
rows (2000 zero + 20000 nonnegative + 2000 SOC blocks of dim 3 + 10 SOC
blocks of dim 50); P = diag(U[0.1,1]) + tridiagonal N(0,0.05) off-diagonals
with the diagonal-dominance repair d_i <- max(d_i, |e_{i-1}|+|e_i|+0.01);
A with 5 nonzeros per row; seeded feasible construction of q and b.  The
problem is set up once (symbolic analysis, untimed); a step is one complete
cold-started solve() to the default 1e-8 tolerances (warm start off, so no
step reuses the previous optimum), on data already resident in HBM.

Prints one JSON line.  Extra fields: "roofline" for the kernel class with the
largest device time inside the timed region (HIP events on the solver's
stream), and "cpu_baseline": the CPU oracle (single-threaded restatement of
the reference method) timed on a bounded sample of the same workload family.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# HIP events around the launches of every 8th factorization and KKT solve: an
# event pair around every launch (about a thousand per factorization) adds
# 6% to a config-1 solve, every 4th about 2%.  Kernel times and the roofline
# cover the factorizations that ran to completion (a checked factorization
# that stops at a failing pivot is left out with its time and its work).
os.environ.setdefault("CONICXX_KERNEL_TIMING", "8")

HBM_PEAK_GBS = 8000.0
# k_trail HBM bytes per launch on config 1 (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0
# per the copy calibration, averaged over the launches of setup + one solve
# that did work: launches of a factorization stopped at a failing pivot
# return at once and are left out, as in the timing)
K_TRAIL_C1_TRAFFIC_BYTES = 3.17e8
FP64_MFMA_PEAK_TFLOPS = 78.6
MI355X_CUS = 256  # compute units: one k_batch_solve workgroup is resident on each


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(full_flops, scale=0.1):
    """Oracle solve of the same generator at `scale`, extrapolated to the full
    instance by the ratio of per-factorization flops (factorization-dominated)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    from conicxx_amd import analyze_pattern
    from conicxx_amd import problems as PR

    p = PR.config1(scale=scale)
    s = oracle_py.OracleSolver()
    assert s.setup(p.P, p.q, p.A, p.b, p.zero_dim, p.nonneg_dim, p.soc_dims, warm_start=0)
    t0 = time.perf_counter()
    r = s.solve()
    t = time.perf_counter() - t0
    _, _, st = analyze_pattern(PR.kkt_lower_pattern(p))
    ratio = full_flops / st["flops"]
    est = t * ratio
    return {
        "value": 1.0 / est, "unit": "QPs/s", "cores": 1, "kind": "port",
        "sample": (f"oracle solve of the config-1 generator at scale {scale} (n={p.n}, m={p.m}, "
                   f"status {r['status']}, {r['iterations']} iterations) took {t:.2f} s on one core; "
                   f"extrapolated to the full instance by the factorization-flop ratio "
                   f"{ratio:.0f} -> {est:.0f} s per solve"),
        "cpu_model": cpu_model(), "nproc": os.cpu_count(),
    }


def reduced_order(p):
    """Order of the batch's reduced KKT matrix for problem p: the variables plus
    every row except the bound rows (nonnegative rows with a single entry),
    which the batched solver folds into the variable block."""
    cnt = np.diff(p.A.tocsr().indptr)
    lo, hi = p.zero_dim, p.zero_dim + p.nonneg_dim
    return p.n + p.m - int((cnt[lo:hi] == 1).sum())


def batch_work(probs):
    """Per-problem algorithmic flops of one dense LDL^T of the reduced matrix
    (order r: sum over columns of c(c-1) with c = r..1, i.e. (r^3 - r)/3) and of
    one reduced solve with its refinement residual (forward and backward
    substitution over the packed triangle 2r(r-1) + r, plus the residual
    rhs - K x: 2 flops per stored entry of P (both triangles), A (twice: A x
    and A'y) and the cone blocks)."""
    r = np.array([reduced_order(p) for p in probs], dtype=np.float64)
    f_fact = (r ** 3 - r) / 3.0
    cone = np.array([p.nonneg_dim + sum(d * d for d in p.soc_dims) for p in probs], dtype=np.float64)
    nnz = np.array([p.P_full().nnz + 2 * p.A.nnz for p in probs], dtype=np.float64)
    f_solve = 2.0 * r * (r - 1.0) + r + 2.0 * (nnz + cone)
    return r, f_fact, f_solve


def oracle_contact_qps(probs, workers):
    """The oracle (single-threaded restatement of the reference method) over
    the given problems: a loop on one core, or a process pool."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    t0 = time.perf_counter()
    if workers <= 1:
        st = [_oracle_solve(p) for p in probs]
    else:
        import multiprocessing as mp
        with mp.get_context("fork").Pool(workers) as pool:
            st = pool.map(_oracle_solve, probs, chunksize=8)
    t = time.perf_counter() - t0
    return sum(1 for v in st if v == 1) / t, t, st


def _oracle_solve(p):
    import oracle_py
    r = oracle_py.solve(p.P, p.q, p.A, p.b, p.zero_dim, p.nonneg_dim, p.soc_dims, warm_start=0)
    return r["status"]


def batched_cpu_baseline(K=4096, cpu_sample=256):
    """The oracle on the host over the first timestep's contact problems: a
    single-core loop over a sample and a process pool over the whole batch.
    Runs before the process touches the GPU (the pool forks)."""
    from conicxx_amd import problems as PR
    p1 = []
    for i in range(K):
        p = PR.config2_problem(i)
        dq, db = PR.config2_perturbation(i, 1, p.n, p.m)
        p1.append(PR.Problem(p.P, p.q + dq, p.A, p.b + db, p.zero_dim, p.nonneg_dim, p.soc_dims))
    cpu1, cpu1_s, _ = oracle_contact_qps(p1[:cpu_sample], 1)
    workers = max(1, min(16, os.cpu_count() or 1))
    cpuN, cpuN_s, ost = oracle_contact_qps(p1, workers)
    return {"value": cpu1, "unit": "QPs/s", "cores": 1, "kind": "port",
            "sample": f"oracle solves of the first {cpu_sample} timestep-1 problems on one core ({cpu1_s:.1f} s)",
            "all_cores": {"value": cpuN, "cores": workers,
                          "sample": f"all {K} timestep-1 problems over a pool of {workers} processes ({cpuN_s:.1f} s)",
                          "non_solved": [i for i, v in enumerate(ost) if v != 1]},
            "cpu_model": cpu_model(), "nproc": os.cpu_count()}


def batched_form(K=4096, timesteps=100, plain_sample=512, cpu=None):
    """The 4096-problem contact batch: update_data + solve per timestep on the
    batched engine (T timesteps of the workload definition), against a
    per-problem loop of the plain GPU solve timed on a sample of the same
    timestep's problems (solve() only, setup untimed), and the oracle on the
    host (single core on a sample, and a process pool over the whole batch)."""
    from conicxx_amd import BatchSolver, ConeSpec, Settings, Solver, Status
    from conicxx_amd import problems as PR

    t0 = time.perf_counter()
    probs = [PR.config2_problem(i) for i in range(K)]
    gen_s = time.perf_counter() - t0
    r, f_fact, f_solve = batch_work(probs)
    st = Settings(warm_start=False)
    b = BatchSolver()
    assert b.setup([(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims)) for p in probs], st)
    b.solve()  # warm-up: graph capture
    t0 = time.perf_counter()
    ncand = b.autotune(1)  # launch configuration timed on this batch (untimed here)
    tune_s = time.perf_counter() - t0
    print(f"[bench] batch autotune: {ncand} candidates in {tune_s:.1f} s -> {b.tuning()}", file=sys.stderr, flush=True)
    b.solve()  # graph capture with the chosen configuration
    solved, wall, dev_ms, flops, nonsolved, wg_us = 0, 0.0, [], 0.0, [], 0.0
    last = None
    for t in range(1, timesteps + 1):
        qs, bs = [], []
        for i, p in enumerate(probs):
            dq, db = PR.config2_perturbation(i, t, p.n, p.m)
            qs.append(p.q + dq)
            bs.append(p.b + db)
        b.update_data(qs, bs)
        t1 = time.perf_counter()
        stt = b.solve()
        wall += time.perf_counter() - t1
        dev_ms.append(b.last_solve_ms())
        ns = int((stt == int(Status.Solved)).sum())
        solved += ns
        nonsolved.append(K - ns)
        stats = [b.stats(k) for k in range(K)]
        cnt = np.array([[s["factorizations"], s["solves"]] for s in stats], dtype=np.float64)
        flops += float((cnt[:, 0] * f_fact + cnt[:, 1] * f_solve).sum())
        wg_us += sum(s["us"] for s in stats)
        if t == 1:
            last = (stt.copy(), qs, bs, cnt)
        if t == 1 or t % 10 == 0:
            print(f"[bench] batch timestep {t}: {ns}/{K} Solved, {dev_ms[-1]:.0f} ms", file=sys.stderr, flush=True)
    qps = solved / wall
    dev_s = sum(dev_ms) / 1e3
    achieved = flops / dev_s / 1e12
    # compulsory HBM bytes per timestep: P, A (values + indices), q, b in; x, s, z out
    hbm = sum(12.0 * (p.P.nnz + p.A.nnz) + 8.0 * (p.n + p.m) + 8.0 * (p.n + 2 * p.m) for p in probs)
    # per-problem loop of the plain GPU solver on the first timestep's problems
    st1, qs1, bs1, cnt1 = last
    loop_s, loop_solved, mism = 0.0, 0, []
    s = Solver()
    for i in range(plain_sample):
        p = probs[i]
        assert s.setup(p.P, qs1[i], p.A, bs1[i], ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims), st)
        t1 = time.perf_counter()
        res = s.solve()
        loop_s += time.perf_counter() - t1
        loop_solved += res.status == Status.Solved
        if int(res.status) != int(st1[i]):
            mism.append(i)
    loop_qps = loop_solved / loop_s
    return {
        "workload": f"config2: {K} friction-contact QPs (c~U{{4..64}}, n=3c, 2n box rows + c SOC dim 3), "
                    f"{timesteps} update_data()+solve() timesteps, cold-started",
        "qps": qps, "unit": "QPs/s", "ms_per_timestep": 1e3 * wall / timesteps,
        "device_ms_per_timestep": sum(dev_ms) / len(dev_ms),
        "solved_fraction": solved / (K * timesteps),
        "non_solved_per_timestep": {"min": min(nonsolved), "max": max(nonsolved),
                                    "mean": sum(nonsolved) / len(nonsolved)},
        "plain_loop_qps": loop_qps,
        "plain_loop_sample": f"the first {plain_sample} problems of timestep 1, one plain Solver, solve() timed",
        "status_mismatches_vs_plain": mism,
        "speedup_vs_plain_loop": qps / loop_qps, "tuning": b.tuning(), "autotune_candidates": ncand,
        "roofline": {"bound": "mfma", "kernel": "k_batch_solve (persistent, FP64)",
                     "achieved": achieved, "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_MFMA_PEAK_TFLOPS,
                     "work": "sum over problems and timesteps of factorizations x (r^3-r)/3 + reduced solves x "
                             "(2r(r-1)+r + refinement residual), r = reduced order (mean "
                             f"{float(r.mean()):.0f}, max {int(r.max())}); device time from HIP events around "
                             "each timestep's graph",
                     "flops_per_timestep": flops / timesteps,
                     "compulsory_hbm_gbs": hbm / (dev_s / timesteps) / 1e9,
                     # the same work over the workgroups' own device clocks (the time each
                     # workgroup spent on its problems): the rate of one workgroup, and of
                     # all CUs while each has a problem (no idle tail)
                     "per_workgroup_gflops": flops / (wg_us / 1e6) / 1e9,
                     "all_cus_busy_tflops": flops / (wg_us / 1e6) * MI355X_CUS / 1e12,
                     "all_cus_busy_frac": flops / (wg_us / 1e6) * MI355X_CUS / 1e12 / FP64_MFMA_PEAK_TFLOPS,
                     "busy_fraction_of_device_time": wg_us / 1e6 / MI355X_CUS / dev_s},
        "cpu_baseline": cpu,
        "autotune_s": tune_s, "generate_s": gen_s,
    }


def differentiation_form(K=1024):
    """The 1024-problem differentiation batch: batched forward (with the adjoint
    factorization) + batched backward of L = w'x; microseconds per problem."""
    from conicxx_amd import BatchSolver, ConeSpec, Settings, Status
    from conicxx_amd import problems as PR

    probs = [PR.config3_problem(i) for i in range(K)]
    st = Settings(warm_start=False)
    b = BatchSolver()
    assert b.setup([(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims)) for p in probs], st)
    b.set_differentiable(True)
    b.solve()  # warm-up
    ws = [p.meta["w"] for p in probs]
    b.backward(ws)
    t0 = time.perf_counter()
    stt = b.solve()
    fwd_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.backward(ws)
    bwd_s = time.perf_counter() - t0
    # per problem, on its workgroup: the backward pass against one forward KKT
    # solve (triangular solves + refinement residuals / KKT solves)
    stats = [b.stats(k) for k in range(K)]
    ok = [s for s in stats if s["us_backward"] > 0 and s["kkt_solves"] > 0]
    mean = lambda f: sum(f(s) for s in ok) / max(1, len(ok))
    bwd_us = mean(lambda s: s["us_backward"])
    adj_us = mean(lambda s: s["us_adjoint_factor"])
    kkt_us = mean(lambda s: (s["us_solve"] + s["us_residual"]) / s["kkt_solves"])
    last_us = mean(lambda s: s["us_last_kkt_solve"])
    # one forward IPM iteration's linear algebra: its factorization (the forward
    # factorizations exclude the adjoint one) plus its KKT solves
    fwd_fact_us = mean(lambda s: (s["us_factor"] - s["us_adjoint_factor"]) / max(1, s["iterations"]))
    it_us = mean(lambda s: (s["us_factor"] - s["us_adjoint_factor"] + s["us_solve"] + s["us_residual"])
                 / max(1, s["iterations"]))
    return {
        "workload": f"config3: {K} QPs n=64, m=112 (zero 16, nonneg 64, 8 SOC dim 4), loss w'x",
        "fwd_bwd_us_per_problem": 1e6 * (fwd_s + bwd_s) / K,
        "fwd_us_per_problem": 1e6 * fwd_s / K, "bwd_us_per_problem": 1e6 * bwd_s / K,
        "device_fwd_ms": b.last_solve_ms(), "device_bwd_ms": b.last_backward_ms(),
        "solved_fraction": float((stt == int(Status.Solved)).mean()),
        "workgroup_backward_us": bwd_us, "workgroup_forward_kkt_solve_us": kkt_us,
        "backward_over_forward_kkt_solve": bwd_us / max(1e-9, kkt_us),
        "workgroup_final_forward_kkt_solve_us": last_us,
        "backward_over_final_forward_kkt_solve": bwd_us / max(1e-9, last_us),
        "workgroup_adjoint_factorization_us": adj_us,
        "workgroup_forward_factorization_us_per_iteration": fwd_fact_us,
        "workgroup_forward_iteration_linear_algebra_us": it_us,
        "backward_with_adjoint_factorization_over_forward_kkt_solve": (bwd_us + adj_us) / max(1e-9, kkt_us),
        "backward_with_adjoint_factorization_over_forward_iteration": (bwd_us + adj_us) / max(1e-9, it_us),
        "backward_reduced_solves": mean(lambda s: s["backward_solves"]),
        "note": "backward = one adjoint KKT solve (with refinement) per problem plus the gradient assembly, "
                "on the adjoint factor (final NT scaling with the SOC tangent correction) that the forward "
                "pass factors after its last iteration; that factorization is the backward's cost and is "
                "reported with it ('..._with_adjoint_factorization_...'): the correction has rank up to "
                "sum(d-2) over the SOC blocks, and refinement on the IPM's own last factor does not "
                "contract on it (spectral radius above 1 on some of these problems). Per-problem times "
                "from device clocks; the forward KKT solve is the mean over the solve (triangular solves + "
                "refinement residuals), the final iteration's solve, and one IPM iteration's "
                "factorization + KKT solves",
    }


FP32_MFMA_PEAK_TFLOPS = 157.3
FACTOR_KINDS = ("trail", "panel", "assemble", "small")


def precision_form(scale=1.0, accuracy_scale=0.1):
    """Precision form on the config-4 generator (n=100000, KKT order N=245000 at
    scale 1): device time of one numeric factorization in FP32 (f32 MFMA
    trailing updates) against FP64 (f64 MFMA) on the same KKT matrices (the
    initial point and the first IPM iteration), and the full-solve error of
    FP32 + FP64 GMRES refinement against the FP64 solve on the same generator
    at `accuracy_scale`."""
    import numpy as np
    from conicxx_amd import ConeSpec, FactorPrecision, Settings, Solver, Status
    from conicxx_amd import problems as PR

    p = PR.config4(scale=scale)
    cs = ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims)
    s = Solver()
    t0 = time.perf_counter()
    assert s.setup(p.P, p.q, p.A, p.b, cs, Settings(max_iter=1, warm_start=False)), "setup failed"
    setup_s = time.perf_counter() - t0
    kkt = s.kkt_stats()
    res = {}
    for name, prec in (("fp64", FactorPrecision.FP64), ("fp32", FactorPrecision.FP32Refined)):
        s.setSettings(Settings(max_iter=1, warm_start=False, factor_precision=prec))
        s.solve()
        ks, pr = s.kernel_stats(), s.profile()
        nf = max(1, int(pr["factorizations"]))
        tr = ks.get("trail", {"ms": 0.0, "work": 0.0})
        res[name] = {"factor_device_ms": sum(ks[k]["ms"] for k in FACTOR_KINDS if k in ks) / nf,
                     "factorizations": nf, "fp32_factorizations": int(pr["fp32_factorizations"]),
                     "trail_tflops": tr["work"] / max(1e-9, tr["ms"]) / 1e9,
                     "per_factorization_ms": {k: ks[k]["ms"] / nf for k in FACTOR_KINDS if k in ks}}
        print(f"[bench] precision {name}: {res[name]['factor_device_ms']:.0f} ms per factorization", file=sys.stderr, flush=True)
    # accuracy of the whole method: full solves, FP32 + refinement vs FP64, same
    # problem (the timing solver's setup is reused at the same scale)
    if accuracy_scale == scale:
        q, sv = p, s
    else:
        del s
        q = PR.config4(scale=accuracy_scale)
        sv = Solver()
        assert sv.setup(q.P, q.q, q.A, q.b, ConeSpec(q.zero_dim, q.nonneg_dim, q.soc_dims),
                        Settings(warm_start=False))
    sols, profs, cert = {}, {}, {}
    for name, prec in (("fp64", FactorPrecision.FP64), ("fp32", FactorPrecision.FP32Refined)):
        sv.setSettings(Settings(warm_start=False, factor_precision=prec))
        sols[name] = sv.solve()
        profs[name] = sv.profile()
        r = sols[name]
        c = PR.kkt_certificate(q, r.x, r.s, r.z)
        sc = 1.0 + float(np.abs(q.q).max() + np.abs(q.b).max())
        cert[name] = {"stationarity_rel": c["stationarity"] / sc, "feasibility_rel": c["feasibility"] / sc,
                      "complementarity": c["complementarity"], "margin_s": c["margin_s"], "margin_z": c["margin_z"]}
        print(f"[bench] precision full solve {name}: {r.status.name} in {r.info.iterations} iterations, "
              f"{r.info.solve_time_s:.2f} s", file=sys.stderr, flush=True)
    a, b = sols["fp64"], sols["fp32"]
    rel = lambda u, v: float(np.linalg.norm(u - v) / max(1e-300, np.linalg.norm(v)))
    return {
        "workload": (f"config4 at scale {scale}: n={p.n}, m={p.m}, KKT N={kkt['N']:.0f}, nnz(L)={kkt['nnz_L']:.3g}, "
                     f"factor flops={kkt['flops']:.3g}, largest front={kkt['max_front']:.0f}; setup {setup_s:.1f} s"),
        "fp64": res["fp64"], "fp32": res["fp32"],
        "factorization_speedup": res["fp64"]["factor_device_ms"] / max(1e-9, res["fp32"]["factor_device_ms"]),
        "trail_roofline": {"fp64_frac": res["fp64"]["trail_tflops"] / FP64_MFMA_PEAK_TFLOPS,
                           "fp32_frac": res["fp32"]["trail_tflops"] / FP32_MFMA_PEAK_TFLOPS},
        "accuracy": {
            "workload": f"config4 at scale {accuracy_scale}: n={q.n}, m={q.m}, full solves to the default 1e-8 tolerances",
            "status": {"fp64": a.status.name, "fp32": b.status.name},
            "iterations": {"fp64": a.info.iterations, "fp32": b.info.iterations},
            "fp32_factorizations": int(profs["fp32"]["fp32_factorizations"]),
            "fp64_factorizations_in_fp32_run": int(profs["fp32"]["factorizations"] - profs["fp32"]["fp32_factorizations"]),
            "fp32_fallback_iteration": int(profs["fp32"]["fallback_iteration"]),
            "rel_err_vs_fp64": {"x": rel(b.x, a.x), "s": rel(b.s, a.s), "z": rel(b.z, a.z),
                                "objective": abs(b.objective - a.objective) / max(1e-300, abs(a.objective))},
            "solve_s": {"fp64": a.info.solve_time_s, "fp32": b.info.solve_time_s},
            "kkt_certificate": cert,
            "error_note": "the FP64 oracle cannot solve this instance in a bounded time on the host; the "
                          "FP32 result's error is stated against the GPU FP64 solve of the same problem, "
                          "and each solve carries its own KKT certificate recomputed on the host",
        },
    }


def copy_calibration(nbytes=1 << 30, reps=3):
    """Three plain device-to-device copies of 1 GiB through the HIP runtime (its
    own copy kernel; no torch kernels, which do not survive counter collection
    here): in a rocprofv3 --pmc run their FETCH_SIZE / WRITE_SIZE per dispatch
    give the scale of the counters against the known 1 GiB moved."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--timesteps", type=int, default=100,
                    help="update_data()+solve() timesteps of the batched form (the workload defines 100)")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the batched, differentiation and precision forms")
    ap.add_argument("--form", choices=["batched", "differentiation", "precision"],
                    help="run only this form and print its JSON object (for profiling)")
    ap.add_argument("--precision-accuracy-scale", type=float, default=0.1,
                    help="generator scale of the precision form's full FP64 and FP32 solves (1.0: the "
                         "full n=100000 instance, about 3.5 minutes)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with three 1 GiB device-to-device copies: the reference kernel that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters of a profiled run")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the plain solve of one QP runs on a single device "
              "(run without --gpus or with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.form:
        if args.form == "batched":
            cpu = None if args.no_cpu_baseline else batched_cpu_baseline()
            res = batched_form(timesteps=args.timesteps, cpu=cpu)
        else:
            res = (differentiation_form() if args.form == "differentiation"
                   else precision_form(accuracy_scale=args.precision_accuracy_scale))
        print(json.dumps({"form": args.form, **res}))
        return
    # the host-only oracle baseline of the batched form forks a process pool:
    # it runs before this process touches the GPU
    batch_cpu = None if (args.no_forms or args.no_cpu_baseline) else batched_cpu_baseline()

    import numpy as np  # noqa: F401
    from conicxx_amd import ConeSpec, Settings, Solver, Status
    from conicxx_amd import problems as PR

    p = PR.config1()
    solver = Solver()
    t0 = time.perf_counter()
    ok = solver.setup(p.P, p.q, p.A, p.b, ConeSpec(p.zero_dim, p.nonneg_dim, p.soc_dims),
                      Settings(warm_start=False))
    setup_s = time.perf_counter() - t0
    assert ok, "setup failed"
    kkt = solver.kkt_stats()
    # exact-structure work of one factorization (sum over columns of c_j (c_j - 1)
    # with the true column counts of the final order): the algorithmic flops;
    # kkt["flops"] counts the stored supernodal structure, explicit zeros of
    # relaxed amalgamation included
    from conicxx_amd import analyze_pattern
    _, cc, _ = analyze_pattern(PR.kkt_lower_pattern(p))
    cc = cc.astype(np.float64)
    flops_exact = float((cc * (cc - 1.0)).sum())
    print(f"[bench] setup {setup_s:.2f} s, N={kkt['N']:.0f}, nnz(L)={kkt['nnz_L']:.3g}, "
          f"factor flops={kkt['flops']:.3g}, largest front={kkt['max_front']:.0f}", file=sys.stderr, flush=True)
    for w in range(args.warmup):
        sol = solver.solve()
        print(f"[bench] warmup {w}: {sol.status.name} in {sol.info.iterations} iterations, "
              f"{sol.info.solve_time_s:.3f} s", file=sys.stderr, flush=True)
    agg = {}
    statuses, iters = [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        sol = solver.solve()
        statuses.append(sol.status)
        iters.append(sol.info.iterations)
        for name, v in solver.kernel_stats().items():
            a = agg.setdefault(name, {"ms": 0.0, "launches": 0, "work": 0.0})
            a["ms"] += v["ms"]
            a["launches"] += v["launches"]
            a["work"] += v["work"]
        print(f"[bench] step {k}: {sol.status.name} in {sol.info.iterations} iterations, "
              f"{sol.info.solve_time_s:.3f} s", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    valid = all(s == Status.Solved for s in statuses)
    value = args.steps / elapsed

    # roofline of the dominant instrumented kernel class
    dom = max(agg, key=lambda k: agg[k]["ms"]) if agg else None
    roof = None
    if dom:
        a = agg[dom]
        avg_s = a["ms"] / max(1, a["launches"]) / 1e3
        per_launch = a["work"] / max(1, a["launches"])
        if dom == "trail":
            # per-launch work counted on the stored structure, scaled to the
            # exact structure by the whole factorization's ratio
            ach_stored = per_launch / avg_s / 1e12
            ach = ach_stored * flops_exact / kkt["flops"]
            roof = {"kernel": "k_trail (FP64 MFMA trailing update)", "bound": "mfma",
                    "achieved": ach, "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": ach / FP64_MFMA_PEAK_TFLOPS,
                    "work": "exact-structure flops (true column counts), per launch: the stored-structure "
                            f"count x {flops_exact / kkt['flops']:.4f} (exact / stored over the factorization)",
                    "achieved_stored_structure": ach_stored,
                    "frac_stored_structure": ach_stored / FP64_MFMA_PEAK_TFLOPS,
                    # a constant: HBM bytes per launch on this workload from the FETCH_SIZE /
                    # WRITE_SIZE counter passes, scaled by a 1 GiB copy, in the profile
                    # summary k_trail_config1_traffic.md (not measured by this run)
                    "traffic": K_TRAIL_C1_TRAFFIC_BYTES if p.n == 20000 else None,
                    "traffic_unit": "bytes per launch (counter profile of this workload; not measured in this run)"}
        else:
            ach = per_launch / avg_s / 1e9 if per_launch else 0.0
            roof = {"kernel": f"k_{dom}", "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None}
        roof["avg_launch_ms"] = a["ms"] / max(1, a["launches"])
        roof["launches"] = a["launches"]
        roof["share_of_device_time"] = a["ms"] / max(1e-9, sum(v["ms"] for v in agg.values()))
    kernel_ms = {k: round(v["ms"] / args.steps, 2) for k, v in agg.items()}
    trail = agg.get("trail")
    trail_tflops = (trail["work"] / (trail["ms"] / 1e3) / 1e12) if trail and trail["ms"] > 0 else None

    out = {
        "metric": "QPs solved/sec at 1e-8 rel gap, batched vs plain GPU; fwd+bwd us per problem",
        "value": value if valid else 0.0,
        "unit": "QPs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, feasible by construction)",
        "config": {"workload": "value: plain FP64 solve of config1, one sparse conic QP, n=20000, m=28500 "
                               "(zero 2000, nonneg 20000, 2000 SOC dim 3, 10 SOC dim 50), cold-started solve "
                               "per step; the batched (config2) and fwd+bwd (config3) figures are in "
                               "'batched' and 'differentiation'",
                   "psd_repair": p.meta.get("psd_repair"), "kkt_N": kkt["N"], "nnz_L": kkt["nnz_L"],
                   "factor_flops": kkt["flops"], "factor_flops_exact": flops_exact,
                   "largest_front": kkt["max_front"],
                   "iterations": iters, "all_solved": valid},
        "roofline": roof,
        "kernel_ms_per_step": kernel_ms,
        "trail_tflops": trail_tflops,
    }
    if not args.no_forms:
        out["batched"] = batched_form(timesteps=args.timesteps, cpu=batch_cpu)
        out["differentiation"] = differentiation_form()
        out["precision"] = precision_form(accuracy_scale=args.precision_accuracy_scale)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(kkt["flops"])
    print(json.dumps(out))
    if args.copy_calibration:
        copy_calibration()


if __name__ == "__main__":
    main()
