"""The benchmark workloads: seeded configurations and request streams.

* single systems: N=500 phi=0.50 (the plain FP64 case), N=4000 phi=0.56
  (rate control, mixed precision), N=32000 phi=0.58 (stress control);
* batched sweep: 64 independent systems, N log-uniform in [250, 2000],
  phi uniform in [0.45, 0.58], shear rate log-uniform over 3 decades in
  repulsive-force units;
* serving: Poisson-arriving requests, N log-uniform in [200, 8000], phi
  uniform in [0.50, 0.58], each running a given strain.

All bidisperse (radii 1.0 / 1.4, 50:50 by volume) and generated by
lf_dem_amd.confgen.
"""
import os

import numpy as np

from . import confgen

SINGLE = {"c1": (500, 0.50, 1), "c2": (4000, 0.56, 1), "c5": (32000, 0.58, 1)}


def sweep_specs(count=64, seed=1):
    """Batched sweep: (n, phi, seed, repulsion) per instance.  The shear
    rate is 10^U(-2, 1) in repulsive units, i.e. repulsion = 1 / rate in
    hydrodynamic units; instance k is seeded seed + k."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(round(np.exp(rng.uniform(np.log(250), np.log(2000)))))
        phi = float(rng.uniform(0.45, 0.58))
        rate_r = float(10 ** rng.uniform(-2, 1))
        out.append((n, phi, seed + k, 1.0 / rate_r))
    return out


def serving_requests(seed, rate, duration, strain):
    """Poisson arrivals (mean `rate` per second) over `duration` seconds:
    (arrival s, n, phi, seed, strain) with N log-uniform in [200, 8000] and
    phi uniform in [0.50, 0.58]."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        n = int(round(np.exp(rng.uniform(np.log(200), np.log(8000)))))
        phi = float(rng.uniform(0.50, 0.58))
        out.append((t, n, phi, seed * 1000 + len(out), strain))
    return out


def generate(n, phi, seed):
    """confgen.generate, kept as conf_<n>_<phi>_<seed>.npz in the directory
    named by LFD_CONF_CACHE when that is set (several runs of one job
    generating the same large configuration)."""
    cache = os.environ.get("LFD_CONF_CACHE")
    if not cache:
        return confgen.generate(n, phi, seed)
    path = os.path.join(cache, "conf_%d_%.4f_%d.npz" % (n, phi, seed))
    if os.path.exists(path):
        d = np.load(path)
        return d["pos"], d["radius"], d["L"]
    pos, rad, L = confgen.generate(n, phi, seed)
    os.makedirs(cache, exist_ok=True)
    tmp = path + ".tmp.npz"
    np.savez(tmp, pos=pos, radius=rad, L=L)
    os.replace(tmp, path)
    return pos, rad, L


def _gen(spec):
    n, phi, seed = spec[:3]
    return confgen.generate(n, phi, seed)


def generate_many(specs, workers=8):
    """Configurations of (n, phi, seed, ...) specs, in a process pool."""
    from concurrent.futures import ProcessPoolExecutor
    if workers <= 1 or len(specs) <= 1:
        return [_gen(s) for s in specs]
    with ProcessPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(_gen, specs))
