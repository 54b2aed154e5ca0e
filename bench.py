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
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "steps/s at N=4000 φ=0.56 vs plain GPU; p50/p99 step latency; rel.err≤1e-8"
HBM_PEAK_GBS = 8000.0


def solver_bytes_per_iter(n, p, waves=0):
    """Algorithmic bytes of one inner FP32 PCG iteration with the 6x6
    block-Jacobi preconditioner: SpMV with LF_DEM's packed 6x6 blocks applied
    in both directions (18 floats + index per pair block, 15 floats + index
    per diagonal block, the row's own 6-vector), the 6x6 FP32 inverses, and
    the pipelined-PCG vector updates.  With the aggregate preconditioner
    (waves > 0) see solver_bytes_per_iter_agg."""
    spmv = 2 * p * (72 + 4) + n * (60 + 8) + 6 * 4 * n
    precond = (AGG_PACK * 4 * waves if waves else 144 * n) + 2 * 6 * 4 * n
    vec = n * 6 * 4 * (4 + 3) + n * 6 * 4 * 4
    return spmv + precond + vec


def solver_bytes_per_iter_rhs(n, p, nrhs):
    """As solver_bytes_per_iter with the 6x6 block-Jacobi preconditioner for
    nrhs right-hand sides solved together: the blocks and inverses are read
    once per iteration for all of them, the vectors once per right-hand side."""
    spmv = 2 * p * (72 + 4) + n * (60 + 8) + nrhs * 6 * 4 * n
    precond = 144 * n + nrhs * 2 * 6 * 4 * n
    vec = nrhs * (n * 6 * 4 * (4 + 3) + n * 6 * 4 * 4)
    return spmv + precond + vec


def solver_bytes_per_iter_agg(n, p_off, waves, nrhs=1):
    """As solver_bytes_per_iter for the aggregate-preconditioned solvers with
    nrhs right-hand sides: their product is Eisenstat's, A u = r + E u for a
    preconditioned u, so they read only the p_off couplings between different
    waves (both directions) and no diagonal block; the packed FP32 wave
    inverses are read once for every right-hand side, the vectors once per
    right-hand side."""
    spmv = 2 * p_off * (72 + 4) + nrhs * 6 * 4 * n
    precond = AGG_PACK * 4 * waves + nrhs * 2 * 6 * 4 * n
    vec = nrhs * (n * 6 * 4 * (4 + 3) + n * 6 * 4 * 4)
    return spmv + precond + vec


def pairs_between_waves(S):
    """Pairs of a system whose particles sit in different waves of the
    aggregate preconditioner."""
    wave_of = S.solver_slots() // 8
    pr = S.pairs()
    return int(np.count_nonzero(wave_of[pr["p0"]] != wave_of[pr["p1"]]))


def solver_bytes_per_outer(n, p):
    """FP64 residual r = b - R x: packed FP64 blocks both ways, diag, x, b, r."""
    return 2 * p * (144 + 4) + n * (120 + 8) + n * 6 * 8 * 3


def measured_traffic(kernel, workload):
    """HBM bytes per launch of the solver kernel from the committed rocprofv3
    PMC summary of this workload (FETCH_SIZE and WRITE_SIZE, each scaled by
    the same counter on a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "%s_%s_pmc_traffic.txt" % (kernel, workload))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


AGG_PACK = 48 * 49 // 2
# lfd_solve_opts.launch of the persistent solvers (lfd_launch)
LAUNCHES = {"auto": 0, "cooperative": 1, "plain": 2}


def solve_error_at_state(S, opts, capi, tight=1e-13):
    """Relative error of one warm-started velocity solve at the system's
    current state (as in the timed steps: the solve starts from the last
    step's velocities) against the same matrix and right-hand side solved on
    to `tight` relative residual, continuing from that solution."""
    S.compute_velocities(True)
    a = S.state()
    t = type(opts)()
    ctypes.memmove(ctypes.byref(t), ctypes.byref(opts), ctypes.sizeof(opts))
    t.tol = tight
    S.set_opts(t)
    S.compute_velocities(True)
    b = S.state()
    st = S.solve_stats()
    S.set_opts(opts)
    out = {k: float(np.linalg.norm(a[k] - b[k]) / np.linalg.norm(b[k])) for k in ("na_vel", "na_ang_vel")}
    out["vs"] = "the same system solved on to %.0e relative residual (reached %.1e)" % (
        tight, st["res_norm"] / st["rhs_norm"] if st["rhs_norm"] else 0.0)
    return out


def run(args):
    from lf_dem_amd import capi, confgen
    if capi.lib().lfd_device_count() < 1:
        raise SystemExit("no GPU visible")
    from lf_dem_amd import workloads
    t0 = time.time()
    pos, rad, L = workloads.generate(args.n, args.phi, args.seed)
    gen_s = time.time() - t0
    params = capi.default_params()
    opts = capi.default_opts(kind=0, tol=args.tol, launch=LAUNCHES[args.launch], precond=args.precond, agg_waves=args.agg_waves)
    if args.inner_tol is not None:
        opts.inner_tol = args.inner_tol
    tuned = None
    if args.autotune:
        from lf_dem_amd import autotune
        opts, tuned = autotune.tune_system(pos, rad, L, params=params, base=opts, cache_path=None)
        args.precond = opts.precond
    S = capi.System(pos, rad, L, params=params, opts=opts)
    S.first_update()
    S.step(args.warmup)
    waves = S.precond_waves()
    st0 = S.solve_stats()
    S.profile(True)
    t0 = time.perf_counter()
    S.step(args.steps)
    elapsed = time.perf_counter() - t0
    solve_ms, launches = S.profile(False)
    st1 = S.solve_stats()
    npairs = S.npairs
    eager_sps = args.steps / elapsed
    # the same steps replayed from captured step graphs (lfd_system_step_graph),
    # the library's mode for a fixed configuration: one capture per pair-buffer
    # parity before the timed region
    graph_sps = None
    if args.graph:
        S.step(2, graph=True)
        if S.graphs() > 0:
            t0 = time.perf_counter()
            S.step(args.steps, graph=True)
            graph_sps = args.steps / (time.perf_counter() - t0)
    sps = graph_sps if graph_sps else eager_sps
    # per-step latency (each step synchronised)
    lat = []
    for _ in range(args.latency_steps):
        t1 = time.perf_counter()
        S.step(1, graph=graph_sps is not None)
        lat.append((time.perf_counter() - t1) * 1e3)
    lat = np.array(lat) if lat else np.zeros(1)
    long_window = run_long_window(S, args) if args.strain_window > 0 else None
    err_final = solve_error_at_state(S, opts, capi)
    inner = st1["total_inner"] - st0["total_inner"]
    iters_per_solve = inner / max(launches, 1)
    # plain single-instance GPU form on the same configuration
    # (kind 1: host-side scalars, the form the workload is quoted against;
    # kind 3: the same iterates with the scalars on the device, to show how
    # much of the ratio is host round trips)
    plain_sps = plain_dev_sps = None
    if args.plain_steps > 0:
        rates = []
        for kind in (1, 3):
            P = capi.System(pos, rad, L, params=params, opts=capi.default_opts(kind=kind, tol=args.tol, max_inner=20000))
            P.first_update()
            P.step(1)
            t0 = time.perf_counter()
            P.step(args.plain_steps)
            rates.append(args.plain_steps / (time.perf_counter() - t0))
            del P
        plain_sps, plain_dev_sps = rates
    # roofline of the dominant kernel (the persistent solver launch)
    avg_ms = solve_ms / max(launches, 1)
    outer_per_solve = float(max(st1["iters_outer"], 1))     # refinement passes of the last solve
    if waves > 0:
        p_off = pairs_between_waves(S)
        per_iter = solver_bytes_per_iter_agg(args.n, p_off, waves)
    else:
        p_off = None
        per_iter = solver_bytes_per_iter(args.n, npairs)
    alg_bytes = iters_per_solve * per_iter + outer_per_solve * solver_bytes_per_outer(args.n, npairs)
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    kernel = ["k_solve_lds", "k_solve_agg"][args.precond]
    default_workload = (args.n, args.phi, args.seed) == (4000, 0.56, 1)
    traffic = measured_traffic(kernel, "n4000_phi056") if default_workload else None
    out = {
        "metric": METRIC, "value": round(sps, 3), "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 / sps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "N=4000 bidisperse 1.0/1.4 phi=0.56 rate-controlled shear, mixed precision",
                   "N": args.n, "phi": args.phi, "seed": args.seed, "pairs": npairs, "tol": args.tol, "launch": args.launch,
                   "precond": ["block-jacobi", "aggregate"][opts.precond], "regroup_interval": opts.regroup_interval,
                   "inner_tol": opts.inner_tol, "agg_waves": opts.agg_waves or 8,
                   "parallelism": "single GPU"},
        "stepping": "captured step graphs" if graph_sps else "eager",
        "eager_steps_per_s": round(eager_sps, 3),
        "plain_gpu_steps_per_s": round(plain_sps, 3) if plain_sps else None,
        "speedup_vs_plain_gpu": round(sps / plain_sps, 2) if plain_sps else None,
        "plain_gpu_device_scalars_steps_per_s": round(plain_dev_sps, 3) if plain_dev_sps else None,
        "speedup_vs_plain_gpu_device_scalars": round(sps / plain_dev_sps, 2) if plain_dev_sps else None,
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3), "steps": len(lat)},
        "long_window": long_window,
        "solver": {"inner_iters_per_solve": round(iters_per_solve, 1), "solve_launches": launches,
                   "avg_solve_ms": round(avg_ms, 4), "precond_waves": waves, "pairs_between_waves": p_off,
                   "last_solve": {"outer_passes": st1["iters_outer"], "inner_iters": st1["iters_inner"],
                                  "rel_residual": st1["res_norm"] / st1["rhs_norm"] if st1["rhs_norm"] else None}},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": round(traffic) if traffic else None,
                     "algorithmic_bytes_per_launch": round(alg_bytes)},
        "confgen_s": round(gen_s, 2),
        "rel_err_final_state": err_final,
    }
    if tuned:
        out["autotune"] = tuned
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
    from lf_dem_amd import workloads
    n, phi = 32000, 0.58
    t0 = time.time()
    pos, rad, L = workloads.generate(n, phi, args.seed)
    gen_s = time.time() - t0
    print("configuration generated in %.1f s" % gen_s, file=sys.stderr, flush=True)
    params = capi.default_params()
    opts = capi.default_opts(kind=0, tol=args.tol, launch=LAUNCHES[args.launch], precond=args.precond,
                             max_rhs=args.max_rhs, agg_waves=args.agg_waves)
    if args.inner_tol is not None:
        opts.inner_tol = args.inner_tol

    def to_stress_control(S):
        S.compute_stress(False)
        target = S.stress()["total"][2]*2*0.5   # S:E_hat with E_hat = (0, 0, 1/2, 0, 0, 0)
        S.set_control("stress", target)
        return target
    tuned = None
    if args.autotune:
        from lf_dem_amd import autotune
        opts, tuned = autotune.tune_system(pos, rad, L, params=params, base=opts, cache_path=None, warmup=2,
                                           steps=4, control="stress", prepare=to_stress_control)
    S = capi.System(pos, rad, L, params=params, opts=opts)
    S.first_update()
    target = to_stress_control(S)
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
    long_window = run_long_window(S, args) if args.strain_window > 0 else None
    err_final = solve_error_at_state(S, opts, capi)
    sps = args.steps/elapsed
    # roofline of the solver launch (the aggregate solver of a large instance
    # or the block-Jacobi LDS solver, one or two right-hand sides per launch
    # as they ran)
    nrhs = max(1, int(round(4 * args.steps / launches))) if launches else 1
    avg_ms = solve_ms / max(launches, 1)
    iters = (st1["total_inner"] - st0["total_inner"]) / max(launches, 1)
    npairs = S.npairs
    waves = S.precond_waves()
    pair_suffix = "_pair" if nrhs == 2 else ""
    p_off = None
    if waves > 0:
        p_off = pairs_between_waves(S)
        per_iter = solver_bytes_per_iter_agg(n, p_off, waves, nrhs)
        kernel = "k_solve_agg_large" + pair_suffix
    else:
        per_iter = solver_bytes_per_iter_rhs(n, npairs, nrhs)
        kernel = "k_solve_lds" + pair_suffix
    alg = iters * per_iter + max(st1["iters_outer"], 1) * nrhs * solver_bytes_per_outer(n, npairs)
    achieved = alg / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic = measured_traffic(kernel, "n32000_phi058") if args.seed == 1 else None
    out = {
        "metric": "steps/s at N=32000 phi=0.58 stress control (2 RHS per solve)", "value": round(sps, 3),
        "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3/sps, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "N=32000 bidisperse 1.0/1.4 phi=0.58 stress-controlled shear", "N": n, "phi": phi,
                   "seed": args.seed, "pairs": S.npairs, "target_stress": target, "tol": args.tol, "launch": args.launch,
                   "parallelism": "single GPU"},
        "solver_form": {0: "mixed, LDS-staged", 1: "plain FP64", 2: "mixed, HBM-streamed"}[opts.kind],
        "precond_in_use": ("aggregate, %d waves, %d of %d pairs between waves" % (waves, p_off, npairs))
                          if waves > 0 else "6x6 block-Jacobi",
        # the large-instance and block-Jacobi LDS solvers pair the two right-hand sides
        # from the launches that ran: a stress-controlled step solves 4 right-hand sides
        "rhs_per_launch": round(4 * args.steps / launches, 2) if launches else None,
        "shear_rate_end": sr["shear_rate"], "viscosity_end": sr["viscosity"],
        "solver": {"solve_launches": launches, "avg_solve_ms": round(solve_ms/max(launches, 1), 4),
                   "inner_iters_per_solve": round((st1["total_inner"]-st0["total_inner"])/max(launches, 1), 1)},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": round(traffic) if traffic else None, "algorithmic_bytes_per_launch": round(alg)},
        "confgen_s": round(gen_s, 1),
        "long_window": long_window,
        "rel_err_final_state": err_final,
    }
    if tuned:
        out["autotune"] = tuned
    print(json.dumps(out), flush=True)


def run_c3(args):
    """Batched sweep of 64 independent instances stepped together: one set of
    kernels per phase and the solver packed into cooperative launches.  The
    same instances stepped one at a time (each a batch of one) are timed
    beside it."""
    from lf_dem_amd import capi
    if capi.lib().lfd_device_count() < 1:
        raise SystemExit("no GPU visible")
    from lf_dem_amd import workloads
    specs = workloads.sweep_specs(args.instances, args.seed)
    t0 = time.time()
    confs = workloads.generate_many(specs)
    gen_s = time.time() - t0
    print("configurations generated in %.1f s" % gen_s, file=sys.stderr, flush=True)
    opts = capi.default_opts(kind=0, tol=args.tol, launch=LAUNCHES[args.launch], precond=args.precond, agg_waves=args.agg_waves)
    if args.inner_tol is not None:
        opts.inner_tol = args.inner_tol
    params = [capi.default_params(repulsion=s[3]) for s in specs]
    tuned = None
    if args.autotune:
        from lf_dem_amd import autotune
        opts, tuned = autotune.tune_batch(confs, params, base=opts, cache_path=None)
    slots = cells = 0
    for (pos, rad, L), p in zip(confs, params):
        sl, ce, _ = capi.footprint(len(rad), rad, L, p)
        slots += sl
        cells += ce
    B = capi.Batch(slots, cells, len(specs), opts=opts)
    for (pos, rad, L), p in zip(confs, params):
        B.add(pos, rad, L, params=p)
    info = B.info()
    B.step(args.warmup)
    B.sync()
    B.profile(True)
    t0 = time.perf_counter()
    B.step(args.steps)
    B.sync()
    elapsed = time.perf_counter() - t0
    solve_ms, launches = B.profile(False)
    eager_bsps = args.steps / elapsed
    # the same steps replayed from captured step graphs when every solver
    # launch of the batch can be captured (plain launches)
    graph_bsps = None
    if args.graph:
        B.step(2, graph=True)
        B.sync()
        if B.graphs() > 0:
            t0 = time.perf_counter()
            B.step(args.steps, graph=True)
            B.sync()
            graph_bsps = args.steps / (time.perf_counter() - t0)
    lat = []
    for _ in range(args.latency_steps):
        t1 = time.perf_counter()
        B.step(1, graph=graph_bsps is not None)
        B.sync()
        lat.append((time.perf_counter() - t1) * 1e3)
    lat = np.array(lat)
    del B
    # the same instances one at a time
    lone_s = 0.0
    lone_steps = 2
    for (pos, rad, L), p in zip(confs, params):
        S = capi.System(pos, rad, L, params=p, opts=opts)
        S.first_update()
        S.step(1)
        t1 = time.perf_counter()
        S.step(lone_steps)
        lone_s += time.perf_counter() - t1
        del S
    total_n = sum(s[0] for s in specs)
    bsps = graph_bsps if graph_bsps else eager_bsps
    lone_sweep_sps = lone_steps / lone_s          # sweeps of all instances per second, one at a time
    out = {
        "metric": "batched sweep: steps/s of 64 instances stepped together", "value": round(bsps, 3),
        "unit": "batch steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 / bsps, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "64 instances, N log-uniform [250, 2000], phi U(0.45, 0.58), rate 10^U(-2, 1) "
                               "repulsive units", "instances": len(specs), "total_N": total_n,
                   "seed": args.seed, "tol": args.tol, "launch": args.launch, "parallelism": "single GPU"},
        "stepping": "captured step graphs" if graph_bsps else "eager",
        "eager_batch_steps_per_s": round(eager_bsps, 3),
        "instance_steps_per_s": round(bsps * len(specs), 1),
        "particle_steps_per_s": round(bsps * total_n, 0),
        "one_at_a_time_steps_per_s": round(lone_sweep_sps, 3),
        "speedup_vs_one_at_a_time": round(bsps / lone_sweep_sps, 2),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3), "steps": len(lat)},
        "solver": {"launches_per_solve": info["solver_launches"], "solve_launches": launches,
                   "avg_solve_ms": round(solve_ms / max(launches, 1), 4)},
        "device_bytes": info["bytes"],
        "confgen_s": round(gen_s, 1),
    }
    out["config"]["precond"] = ["block-jacobi", "aggregate"][opts.precond]
    if tuned:
        out["autotune"] = tuned
    print(json.dumps(out), flush=True)


def _serve_phase(args, rate, duration, drain, seed, pregen):
    """One serving run: Poisson requests at `rate` for `duration` seconds,
    then up to `drain` seconds to finish the requests that arrived."""
    from lf_dem_amd import capi, confgen, serve, workloads
    arr = workloads.serving_requests(seed, rate, duration, args.strain)
    # the first requests' configurations are the clients' inputs, prepared
    # before the clock starts; later ones are generated at admission
    k = min(len(arr), pregen)
    t0 = time.time()
    confs = workloads.generate_many([(n, phi, sd) for (_, n, phi, sd, _) in arr[:k]], workers=16)
    gen_s = time.time() - t0
    print("%d of %d request configurations generated in %.1f s" % (k, len(arr), gen_s), file=sys.stderr,
          flush=True)
    opts = capi.default_opts(kind=0, tol=args.tol, launch=LAUNCHES[args.launch], precond=args.precond, agg_waves=args.agg_waves)
    if args.inner_tol is not None:
        opts.inner_tol = args.inner_tol
    reqs = []
    for i, a in enumerate(arr):
        if i < k:
            pos, rad, L = confs[i]
            reqs.append(serve.Request(i, pos, rad, L, strain=a[4], arrival=a[0]))
        else:
            n1, n2 = confgen.bidisperse_counts(a[1])
            rad = np.concatenate([np.full(n1, 1.0), np.full(n2, 1.4)])
            L = np.full(3, confgen.box_size(rad, a[2]))
            reqs.append(serve.Request(i, None, rad, L, strain=a[4], arrival=a[0],
                                      loader=lambda a=a: confgen.generate(a[1], a[2], a[3])[0]))
    slot_cap = args.serve_slots
    srv = serve.Server(slot_cap, slot_cap // 4, 1024, opts=opts, keep_state=False, graph=args.graph)
    # warm the kernels on a throwaway request
    w = serve.Request(-1, *confs[0], strain=1e-3)
    srv.submit(w)
    srv.admit(0)
    while srv.live:
        srv.step()
    srv.done.clear()
    srv.step_ms.clear()
    srv.particle_steps = srv.instance_steps = 0
    elapsed = srv.run(reqs, duration=duration, drain=drain)
    summ = srv.summary(elapsed)
    info = srv.batch.info()
    del srv
    r3 = lambda d: None if d is None else {k: round(v, 3) for k, v in d.items()}
    return {"rate_per_s": rate, "arrivals_s": duration, "drain_s": drain, "requests": len(reqs),
            "elapsed_s": round(elapsed, 2), "completed": summ["completed"], "failed": summ["failed"],
            "completed_per_s": round(summ["completed_per_s"], 3), "live_at_end": summ["live_at_end"],
            "waiting_at_end": summ["waiting_at_end"], "response_s": r3(summ["response_s"]),
            "steps_per_request": r3(summ["steps_per_request"]), "admission_wait_s": r3(summ["admission_wait_s"]),
            "step_ms": r3(summ["step_ms"]), "batch_steps": summ["batch_steps"],
            "instance_steps_per_s": round(summ["instance_steps_per_s"], 1),
            "particle_steps_per_s": round(summ["particle_steps_per_s"], 0),
            "slot_cap": info["slot_cap"], "device_bytes": info["bytes"], "confgen_s": round(gen_s, 1),
            "stepping": "captured step graphs, recaptured when the batch changes" if args.graph else "eager"}


def run_c4(args):
    """Serving: Poisson-arriving simulation requests admitted into one batch
    by footprint as room frees up, stepped together, retired at their strain
    (lf_dem_amd.serve).  Two runs: below capacity (--sub-rate requests/s, 0.25, for
    --sub-duration s, then served to completion) for request response times,
    and at the offered load of the workload (--rate, 20/s, for --duration s),
    which exceeds the capacity, for the saturated throughput and batch-step
    latency."""
    from lf_dem_amd import capi
    if capi.lib().lfd_device_count() < 1:
        raise SystemExit("no GPU visible")
    sub = _serve_phase(args, args.sub_rate, args.sub_duration, args.sub_drain, args.seed + 1, 10 ** 6)
    print("below capacity: %s" % json.dumps(sub), file=sys.stderr, flush=True)
    sat = _serve_phase(args, args.rate, args.duration, 0.0, args.seed, args.serve_pregen)
    out = {
        "metric": "serving with continuous batching: completed requests/s and response p50/p99 below capacity; "
                  "particle-steps/s and batch-step p50/p99 at 20 requests/s", "value": sub["completed_per_s"],
        "unit": "requests/s", "n_gpus": 1, "steps": sub["batch_steps"], "warmup": 1,
        "ms_per_step": sub["step_ms"]["p50"] if sub["step_ms"] else None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "Poisson requests, N log-uniform [200, 8000], phi U(0.50, 0.58), %.2f strain each"
                               % args.strain, "seed": args.seed, "tol": args.tol, "launch": args.launch, "parallelism": "single GPU"},
        "latency_ms": sub["step_ms"],
        "response_s": sub["response_s"],
        "below_capacity": sub,
        "saturation": sat,
    }
    print(json.dumps(out), flush=True)
    if sub["completed"] == 0:
        raise SystemExit("serving: no request completed below capacity")


def run_long_window(S, args, chunk=100):
    """The system stepped on (captured step graphs when available) until the
    strain reaches args.strain_window: steps/s, inner iterations per solve and
    pairs over segments of an eighth of that strain, to show the rate once
    contacts have built up (the timed steps are the first few dozen)."""
    graph = S.graphs() > 0
    seg_len = args.strain_window / 8
    st = S.state()
    strain = st["strain"]
    seg_start, seg_steps, seg_t, seg_inner = strain, 0, 0.0, S.solve_stats()["total_inner"]
    total_steps, total_t, segs = 0, 0.0, []
    while strain < args.strain_window - 1e-12:
        t0 = time.perf_counter()
        S.step(chunk, graph=graph)
        S.sync()
        dt = time.perf_counter() - t0
        strain = S.state()["strain"]
        seg_steps += chunk
        seg_t += dt
        total_steps += chunk
        total_t += dt
        if strain >= seg_start + seg_len or strain >= args.strain_window - 1e-12:
            inner = S.solve_stats()["total_inner"]
            segs.append({"strain": round(strain, 4), "steps": seg_steps, "steps_per_s": round(seg_steps / seg_t, 2),
                         "inner_iters_per_solve": round((inner - seg_inner) / (2 * seg_steps), 1),
                         "pairs": S.npairs})
            print("long window: strain %.3f, %d steps, %.1f steps/s" % (strain, seg_steps, seg_steps / seg_t),
                  file=sys.stderr, flush=True)
            seg_start, seg_steps, seg_t, seg_inner = strain, 0, 0.0, inner
    return {"strain_end": round(strain, 4), "steps": total_steps, "steps_per_s": round(total_steps / total_t, 2),
            "stepping": "captured step graphs" if graph else "eager", "segments": segs}


def run_copy(args):
    """Counter calibration: device-to-device copies of 256 MiB (64 Mi float32)
    whose HBM traffic is known, for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE."""
    import torch
    x = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    x.fill_(1.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        y = x.clone()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / 3
    del y
    print(json.dumps({"metric": "device copy bandwidth (counter calibration)", "value": round(2 * x.numel() * 4 / el / 1e9, 1),
                      "unit": "GB/s", "bytes_read": x.numel() * 4, "bytes_written": x.numel() * 4, "copies": 3}),
          flush=True)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _host_cores():
    """Cores this process may use: its affinity set, capped by the thread
    budget the environment sets (OMP_NUM_THREADS)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(n, 1)


def cpu_leg(pos, rad, L, params, opts, args):
    """FP64 direct (oracle) solve at t=0 for the error of the GPU velocities;
    then the CPU baseline: one predictor-corrector step of the oracle (two
    FP64 assemblies and sparse Cholesky factorisations and solves, pair and
    cell updates) on the host, by phase, on 1 core, and the same step run
    concurrently on every core this process may use (independent copies:
    the aggregate rate of one simulation per core)."""
    import subprocess
    import tempfile
    from lf_dem_amd import capi
    from oracle import oracle as O
    S = capi.System(pos, rad, L, params=params, opts=opts)
    S.first_update()
    S.compute_velocities(True)
    g = S.state()
    R = O.System(pos, rad, L)
    R.first_update()
    R.set_in_predictor(True)
    R.compute_velocities()
    r = R.state()
    err = {k: float(np.linalg.norm(g[k] - r[k]) / np.linalg.norm(r[k])) for k in ("na_vel", "na_ang_vel")}
    del S, R
    with tempfile.TemporaryDirectory() as tmp:
        conf = os.path.join(tmp, "conf.npz")
        np.savez(conf, pos=pos, radius=rad, L=L)
        cmd = [sys.executable, "-m", "oracle.cpu_baseline", conf, "1"]

        def run(k):
            procs = [subprocess.Popen(cmd, cwd=ROOT, stdout=subprocess.PIPE, text=True) for _ in range(k)]
            t0 = time.perf_counter()
            outs = [json.loads(p.communicate()[0].strip().splitlines()[-1]) for p in procs]
            if any(p.returncode for p in procs):
                raise SystemExit("cpu baseline: the oracle step failed")
            return outs, time.perf_counter() - t0
        print("cpu baseline: one oracle step on 1 core", file=sys.stderr, flush=True)
        one, _ = run(1)
        cores = _host_cores()
        print("cpu baseline: %d concurrent oracle steps" % cores, file=sys.stderr, flush=True)
        many, wall = run(cores) if cores > 1 else (one, one[0]["step_s"])
    t1 = one[0]["step_s"]
    cpu = {"value": round(1.0 / t1, 5), "unit": "steps/s", "cores": 1, "kind": "port",
           "sample": "one predictor-corrector step of the oracle at N=%d (FP64 assembly, minimum-degree sparse "
                     "Cholesky, two factorisations and solves, %.1f s) from the t=0 state" % (args.n, t1),
           "phases_s": {k: round(v, 4) for k, v in one[0]["phases"].items()},
           "all_cores": {"value": round(cores / wall, 5), "unit": "steps/s", "cores": cores,
                         "sample": "%d independent copies of the same step run concurrently (%.1f s wall)"
                                   % (cores, wall)},
           "cpu_model": _cpu_model(), "nproc": os.cpu_count()}
    return err, cpu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=4000)
    ap.add_argument("--phi", type=float, default=0.56)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--launch", choices=sorted(LAUNCHES), default="auto",
                    help="persistent solver launches: the library default (auto: plain when a launch fits the "
                         "device at once), always cooperative, or always plain")
    ap.add_argument("--tol", type=float, default=1e-10,
                    help="relative residual of the FP64 refinement (the N=4000 workload states 1e-8)")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--latency-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--strain-window", type=float, default=0.0,
                    help="c2, c5: after the timed steps, step on to this strain and report steps/s per segment")
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--rate", type=float, default=20.0, help="c4: request arrivals per second (saturation run)")
    ap.add_argument("--duration", type=float, default=25.0, help="c4: seconds of arrivals and service (saturation)")
    ap.add_argument("--sub-rate", type=float, default=0.25, help="c4: arrivals per second of the run below capacity")
    ap.add_argument("--sub-duration", type=float, default=80.0, help="c4: seconds of arrivals below capacity")
    ap.add_argument("--sub-drain", type=float, default=200.0,
                    help="c4: seconds to finish the requests of the run below capacity after the arrivals")
    ap.add_argument("--strain", type=float, default=0.5, help="c4: strain units per request")
    ap.add_argument("--serve-slots", type=int, default=131072, help="c4: particle slots of the batch")
    ap.add_argument("--serve-pregen", type=int, default=96, help="c4: configurations prepared up front")
    ap.add_argument("--precond", type=int, default=None,
                    help="0: 6x6 block-Jacobi, 1: aggregate (default 1; c3: 0); the autotuner overrides it")
    ap.add_argument("--agg-waves", type=int, default=None, choices=(8, 12, 16),
                    help="aggregate preconditioner: waves of 8 particles per solver workgroup (default: the "
                         "library's, 8; c4: 16); the autotuner overrides it")
    ap.add_argument("--inner-tol", type=float, default=None,
                    help="c5: residual drop per FP32 inner solve (default: the library's)")
    ap.add_argument("--max-rhs", type=int, default=0,
                    help="c5: right-hand sides per launch of the block-Jacobi solver (0: two, 1: one)")
    ap.add_argument("--no-graph", dest="graph", action="store_false",
                    help="c2: time eager steps only (no captured step graphs)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="c2/c3/c5: skip the solver-option autotuner (lf_dem_amd.autotune) before the timed run")
    ap.add_argument("--workload", choices=("c2", "c3", "c4", "c5", "copy"), default="c2",
                    help="c2: N=4000 rate control (headline); c3: batched sweep of 64 instances; "
                         "c4: served requests; c5: N=32000 stress control; copy: 256 MiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.precond is None:
        args.precond = 0 if args.workload == "c3" else 1
    if args.agg_waves is None:
        args.agg_waves = 16 if args.workload == "c4" else 0
    if args.gpus != 1:
        raise SystemExit("bench.py measures one GPU: this build has no multi-GPU path (use --gpus 1)")
    if args.workload == "copy":
        run_copy(args)
    elif args.workload == "c5":
        run_c5(args)
    elif args.workload == "c3":
        run_c3(args)
    elif args.workload == "c4":
        run_c4(args)
    else:
        run(args)


if __name__ == "__main__":
    main()
