"""Benchmark of the MI355X Stokes-resistance step (one GPU).

Workload (the configuration the headline metric is quoted on): N=4000
bidisperse spheres (radii 1.0 / 1.4, 50:50 by volume), phi=0.56, seeded dense
configuration, rate-controlled simple shear with lubrication (delta=1e-3,
cutoff gap 0.5), spring-dashpot contacts (kn=1e4, kt=5e3, mu=1) and
short-range repulsion; FP32 Krylov with FP64 refinement to a 1e-12 relative
residual.  A "step" is one full predictor-corrector step: 2 assemblies,
2 solves, 2 pair updates, 1 cell-list / neighbour rebuild.

Prints one JSON line (rank 0).  Also measured in the same run:
  * the plain single-instance GPU form (FP64 PCG, one kernel per operation),
  * p50 / p99 step latency,
  * relative error of the GPU velocities against the FP64 direct solve,
  * the CPU baseline: the oracle's FP64 sparse-Cholesky solve on the host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "steps/s at N=4000 φ=0.56 vs plain GPU; p50/p99 step latency; rel.err≤1e-8"
HBM_PEAK_GBS = 8000.0


def solver_bytes_per_iter(n, p):
    """Algorithmic bytes of one inner FP32 PCG iteration (SpMV of packed
    6x6 blocks in both directions + block-Jacobi apply + vector updates)."""
    spmv = 2 * p * (72 + 4) + n * (60 + 8) + 6 * 4 * n       # blocks+index both ways, diag, x gather
    vec = n * 6 * 4 * (4 + 3) + n * 144 + n * 6 * 4 * 4      # phase A r/w + M^-1, phase B p/q r/w
    return spmv + vec


def solver_bytes_per_outer(n, p):
    """FP64 residual r = b - R x: packed FP64 blocks both ways, diag, x, b, r."""
    return 2 * p * (144 + 4) + n * (120 + 8) + n * 6 * 8 * 3


def run(args):
    from lf_dem_amd import capi, confgen
    if capi.lib().lfd_device_count() < 1:
        raise SystemExit("no GPU visible")
    t0 = time.time()
    pos, rad, L = confgen.generate(args.n, args.phi, args.seed)
    gen_s = time.time() - t0
    params = capi.default_params()
    opts = capi.default_opts(kind=0, tol=args.tol)
    S = capi.System(pos, rad, L, params=params, opts=opts)
    S.first_update()
    S.step(args.warmup)
    st0 = S.solve_stats()
    S.profile(True)
    t0 = time.perf_counter()
    S.step(args.steps)
    elapsed = time.perf_counter() - t0
    solve_ms, launches = S.profile(False)
    st1 = S.solve_stats()
    npairs = S.npairs
    sps = args.steps / elapsed
    # per-step latency (each step synchronised)
    lat = []
    for _ in range(args.latency_steps):
        t1 = time.perf_counter()
        S.step(1)
        lat.append((time.perf_counter() - t1) * 1e3)
    lat = np.array(lat)
    inner = st1["total_inner"] - st0["total_inner"]
    iters_per_solve = inner / max(launches, 1)
    # plain single-instance GPU form on the same configuration
    P = capi.System(pos, rad, L, params=params, opts=capi.default_opts(kind=1, tol=args.tol, max_inner=20000))
    P.first_update()
    P.step(1)
    t0 = time.perf_counter()
    P.step(args.plain_steps)
    plain_sps = args.plain_steps / (time.perf_counter() - t0)
    del P
    # roofline of the dominant kernel (the persistent solver launch)
    avg_ms = solve_ms / max(launches, 1)
    outer_per_solve = float(max(st1["iters_outer"], 1))     # refinement passes of the last solve
    alg_bytes = iters_per_solve * solver_bytes_per_iter(args.n, npairs) + \
        outer_per_solve * solver_bytes_per_outer(args.n, npairs)
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    out = {
        "metric": METRIC, "value": round(sps, 3), "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 / sps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "N=4000 bidisperse 1.0/1.4 phi=0.56 rate-controlled shear, mixed precision",
                   "N": args.n, "phi": args.phi, "seed": args.seed, "pairs": npairs, "tol": args.tol,
                   "parallelism": "single GPU"},
        "plain_gpu_steps_per_s": round(plain_sps, 3),
        "speedup_vs_plain_gpu": round(sps / plain_sps, 2),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3), "steps": len(lat)},
        "solver": {"inner_iters_per_solve": round(iters_per_solve, 1), "solve_launches": launches,
                   "avg_solve_ms": round(avg_ms, 4)},
        "roofline": {"bound": "hbm", "kernel": "k_solve_lds", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None},
        "confgen_s": round(gen_s, 2),
    }
    if not args.no_cpu:
        out["rel_err"], out["cpu_baseline"] = cpu_leg(pos, rad, L, params, opts, args)
    print(json.dumps(out), flush=True)


def run_c5(args):
    """Stress-controlled near-jamming workload: N=32000 bidisperse, phi=0.58,
    mu=1, fixed shear stress; every velocity computation solves the rate
    independent and rate proportional right-hand sides on one preconditioner
    and sets the shear rate on the device.  The target stress is the shear
    stress of the initial configuration at unit shear rate, so the run starts
    at a rate near 1."""
    from lf_dem_amd import capi, confgen
    if capi.lib().lfd_device_count() < 1:
        raise SystemExit("no GPU visible")
    n, phi = 32000, 0.58
    t0 = time.time()
    pos, rad, L = confgen.generate(n, phi, args.seed)
    gen_s = time.time() - t0
    print("configuration generated in %.1f s" % gen_s, file=sys.stderr, flush=True)
    params = capi.default_params()
    opts = capi.default_opts(kind=0, tol=args.tol)
    S = capi.System(pos, rad, L, params=params, opts=opts)
    S.first_update()
    S.compute_stress(False)
    st = S.stress()
    target = st["total"][2]*2*0.5          # S:E_hat with E_hat = (0, 0, 1/2, 0, 0, 0)
    S.set_control("stress", target)
    S.step(args.warmup)
    st0 = S.solve_stats()
    S.profile(True)
    t0 = time.perf_counter()
    S.step(args.steps)
    elapsed = time.perf_counter() - t0
    solve_ms, launches = S.profile(False)
    st1 = S.solve_stats()
    S.step_stress()
    sr = S.stress()
    sps = args.steps/elapsed
    out = {
        "metric": "steps/s at N=32000 phi=0.58 stress control (2 RHS per solve)", "value": round(sps, 3),
        "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3/sps, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "N=32000 bidisperse 1.0/1.4 phi=0.58 stress-controlled shear", "N": n, "phi": phi,
                   "seed": args.seed, "pairs": S.npairs, "target_stress": target, "tol": args.tol,
                   "parallelism": "single GPU"},
        "shear_rate_end": sr["shear_rate"], "viscosity_end": sr["viscosity"],
        "solver": {"solve_launches": launches, "avg_solve_ms": round(solve_ms/max(launches, 1), 4),
                   "inner_iters_per_solve": round((st1["total_inner"]-st0["total_inner"])/max(launches, 1), 1)},
        "confgen_s": round(gen_s, 1),
    }
    print(json.dumps(out), flush=True)


def cpu_leg(pos, rad, L, params, opts, args):
    """FP64 direct (oracle) solve at t=0: error of the GPU velocities and the
    host-core time of one assembly + sparse Cholesky factorisation + solve."""
    from lf_dem_amd import capi
    from oracle import oracle as O
    S = capi.System(pos, rad, L, params=params, opts=opts)
    S.first_update()
    S.compute_velocities(True)
    g = S.state()
    R = O.System(pos, rad, L)
    R.first_update()
    R.set_in_predictor(True)
    t0 = time.perf_counter()
    R.compute_velocities()
    t_cv = time.perf_counter() - t0
    r = R.state()
    err = {k: float(np.linalg.norm(g[k] - r[k]) / np.linalg.norm(r[k])) for k in ("na_vel", "na_ang_vel")}
    cpu = {"value": round(1.0 / (2 * t_cv), 5), "unit": "steps/s", "cores": 1, "kind": "port",
           "sample": "one FP64 assembly + minimum-degree sparse Cholesky factorisation + solve at N=%d "
                     "(%.1f s); a step is two of these" % (args.n, t_cv)}
    return err, cpu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=4000)
    ap.add_argument("--phi", type=float, default=0.56)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--tol", type=float, default=1e-12)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--latency-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("c2", "c5"), default="c2",
                    help="c2: N=4000 rate control (headline); c5: N=32000 stress control")
    args = ap.parse_args()
    if args.gpus != 1:
        raise SystemExit("bench.py measures one GPU: this build has no multi-GPU path (use --gpus 1)")
    if args.workload == "c5":
        run_c5(args)
    else:
        run(args)


if __name__ == "__main__":
    main()
