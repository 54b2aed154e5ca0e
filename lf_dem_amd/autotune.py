"""Autotuner of the solver options, by measurement on the workload itself.

The right solver form depends on the workload more than on the hardware: for
one instance the aggregate preconditioner (inverses of 48x48 blocks of up to
8 coupled particles) cuts the Krylov iterations by a third; for a large batch
of small instances the 6x6 block-Jacobi solver packs more particles into each
co-resident cooperative launch (16 wavefronts per CU against 12) and needs
fewer launches per solve.  The grouping interval of the aggregate
preconditioner and the inner FP32 tolerance trade set-up cost against
iterations.

`tune_system` and `tune_batch` time each candidate option set on short runs
of the caller's own configuration(s) (every candidate starts from the same
initial state and is timed over the same steps after the same warm-up) and
return the fastest, with the measurements.  Every candidate solves to the same
FP64 residual tolerance, so the choice never trades accuracy for speed.
Results are cached per workload signature (size, volume fraction, number of
instances, control) in a JSON file, so a service tunes once per kind of
workload.

LF_DEM has no counterpart (CHOLMOD's supernodal factorisation has no
options it exposes; StokesSolver.cpp:950-978).
"""
import json
import os
import time

import numpy as np

from . import capi

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "lf_dem_amd", "autotune.json")
# short trials separate candidates only by more than their run-to-run noise:
# among the candidates within MARGIN of the fastest, the one listed first wins
MARGIN = 0.03

# (label, SolveOpts overrides) tried for one instance
SYSTEM_CANDIDATES = [
    ("aggregate, regroup 4", dict(precond=1, regroup_interval=4)),
    ("aggregate, regroup 8", dict(precond=1, regroup_interval=8)),
    ("aggregate, regroup 2", dict(precond=1, regroup_interval=2)),
    ("aggregate, regroup 4, inner 3e-3", dict(precond=1, regroup_interval=4, inner_tol=3e-3)),
    # 12 waves per solver workgroup (inverses in LDS, fewer workgroups)
    ("aggregate, regroup 4, 12 waves per workgroup", dict(precond=1, regroup_interval=4, agg_waves=12)),
    # 16 waves per workgroup (the large-instance form: couplings between waves
    # only, both right-hand sides of stress control in one launch)
    ("aggregate, regroup 4, 16 waves per workgroup", dict(precond=1, regroup_interval=4, agg_waves=16)),
    ("block-jacobi", dict(precond=0)),
    ("block-jacobi, inner 3e-3", dict(precond=0, inner_tol=3e-3)),
    # blocks streamed from HBM every iteration (no LDS staging), with the
    # SpMV lanes per 6x6 block row as the tuned tile parameter
    ("hbm blocks, 2 lanes per row", dict(kind=2, precond=0, lanes_per_row=2)),
    ("hbm blocks, 4 lanes per row", dict(kind=2, precond=0, lanes_per_row=4)),
    ("hbm blocks, 8 lanes per row", dict(kind=2, precond=0, lanes_per_row=8)),
]
# largest instance the aggregate preconditioner groups whole, with the 8- or
# 12-wave solver; above it (LFD_PRECOND_AGGREGATE in include/lfdem_amd.h) the
# instance is grouped in spatial chunks and solved with 16 waves per
# workgroup, up to 16 x 8 particles per compute unit of the device (256 on
# MI355X); a larger one runs 6x6 block-Jacobi whatever is asked
AGGREGATE_WHOLE_MAX_N = 8192
AGGREGATE_MAX_N = 16 * 8 * 256


def system_candidates(n):
    """SYSTEM_CANDIDATES that differ for an instance of n particles: above
    AGGREGATE_WHOLE_MAX_N without the waves-per-workgroup choice (the large
    instances' solver has one form), above AGGREGATE_MAX_N without the
    aggregate ones, which would all time the block-Jacobi solver."""
    if n <= AGGREGATE_WHOLE_MAX_N:
        return list(SYSTEM_CANDIDATES)
    if n <= AGGREGATE_MAX_N:
        return [c for c in SYSTEM_CANDIDATES if "agg_waves" not in c[1]]
    return [c for c in SYSTEM_CANDIDATES if c[1].get("precond", 1) != 1]


# (label, SolveOpts overrides) tried for a batch
BATCH_CANDIDATES = [
    ("block-jacobi", dict(precond=0)),
    ("aggregate, regroup 4", dict(precond=1, regroup_interval=4)),
    ("aggregate, regroup 4, 12 waves per workgroup", dict(precond=1, regroup_interval=4, agg_waves=12)),
    ("aggregate, regroup 4, 16 waves per workgroup", dict(precond=1, regroup_interval=4, agg_waves=16)),
]


def _signature(kind, sizes, phis, control):
    n = np.asarray(sizes, dtype=float)
    return "%s|n=%d|N=%.2g|phi=%.2f|%s" % (kind, len(n), float(n.sum()), float(np.mean(phis)), control)


def _phi(rad, L):
    return float(np.sum(4.0 / 3 * np.pi * np.asarray(rad) ** 3) / np.prod(np.asarray(L)))


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, cache):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def _opts(base, over):
    o = capi.default_opts()
    if base is not None:
        for name, _ in capi.SolveOpts._fields_:
            setattr(o, name, getattr(base, name))
    for k, v in over.items():
        setattr(o, k, v)
    return o


def _time_system(pos, rad, L, params, opts, warmup, steps, prepare=None):
    S = capi.System(pos, rad, L, params=params, opts=opts)
    S.first_update()
    if prepare is not None:
        prepare(S)
    S.step(warmup)
    S.sync()
    t0 = time.perf_counter()
    S.step(steps)
    S.sync()
    return steps / (time.perf_counter() - t0)


def _time_batch(confs, params, opts, warmup, steps):
    slots = cells = 0
    for (pos, rad, L), p in zip(confs, params):
        sl, ce, _ = capi.footprint(len(rad), rad, L, p)
        slots += sl
        cells += ce
    B = capi.Batch(slots, cells, len(confs), opts=opts)
    for (pos, rad, L), p in zip(confs, params):
        B.add(pos, rad, L, params=p)
    B.step(warmup)
    B.sync()
    t0 = time.perf_counter()
    B.step(steps)
    B.sync()
    return steps / (time.perf_counter() - t0)


def _tune(sig, candidates, timer, base, cache_path, use_cache):
    cache = _load(cache_path) if cache_path else {}
    if use_cache and sig in cache:
        hit = cache[sig]
        return _opts(base, hit["options"]), dict(hit, cached=True)
    t0 = time.time()
    trials = []
    for label, over in candidates:
        try:
            rate = timer(_opts(base, over))
        except capi.LfdError as e:          # a form that cannot take this workload
            trials.append({"candidate": label, "steps_per_s": None, "error": str(e)})
            continue
        trials.append({"candidate": label, "steps_per_s": round(rate, 3)})
    ok = [(t["steps_per_s"], i) for i, t in enumerate(trials) if t["steps_per_s"]]
    if not ok:
        raise RuntimeError("autotune: no candidate could run this workload")
    top = max(r for r, _ in ok)
    best = candidates[min(i for r, i in ok if r >= top / (1 + MARGIN))]
    out = {"choice": best[0], "options": best[1], "trials": trials, "seconds": round(time.time() - t0, 2),
           "cached": False}
    if cache_path:
        cache[sig] = {k: out[k] for k in ("choice", "options", "trials")}
        _store(cache_path, cache)
    return _opts(base, best[1]), out


def tune_system(pos, rad, L, params=None, base=None, warmup=3, steps=10, candidates=None,
                cache_path=DEFAULT_CACHE, use_cache=True, control="rate", prepare=None):
    """Fastest SolveOpts for one instance of this configuration: (opts, report).
    prepare(system), when given, runs after the first update of every
    candidate's system (e.g. to switch it to stress control)."""
    params = params if params is not None else capi.default_params()
    sig = _signature("system", [len(rad)], [_phi(rad, L)], control)
    return _tune(sig, candidates or system_candidates(len(rad)),
                 lambda o: _time_system(pos, rad, L, params, o, warmup, steps, prepare), base, cache_path,
                 use_cache)


def tune_batch(confs, params, base=None, warmup=2, steps=6, candidates=None, cache_path=DEFAULT_CACHE,
               use_cache=True, control="rate"):
    """Fastest SolveOpts for a batch of these configurations [(pos, rad, L)]: (opts, report)."""
    sig = _signature("batch", [len(c[1]) for c in confs], [_phi(c[1], c[2]) for c in confs], control)
    return _tune(sig, candidates or BATCH_CANDIDATES, lambda o: _time_batch(confs, params, o, warmup, steps), base,
                 cache_path, use_cache)
