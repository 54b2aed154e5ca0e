"""CPU baseline runner (test infrastructure: only bench.py's cpu_baseline
leg runs it).  Times predictor-corrector steps of the oracle -- the FP64 CPU
restatement of the reference's step path with its sparse Cholesky solve,
one thread -- on a configuration file, split by phase.

    python -m oracle.cpu_baseline conf.npz [nsteps]

conf.npz holds pos (N, 3), radius (N,), L (3,).  Prints one JSON line:
{"steps": n, "step_s": seconds per step, "phases": {phase: seconds per step}}.
"""
import json
import sys
import time

import numpy as np

from oracle import oracle as O

PHASES = ("move", "neighbor", "pair_update", "assemble", "order", "factor", "solve")


def main():
    d = np.load(sys.argv[1])
    nsteps = int(sys.argv[2]) if len(sys.argv) > 2 else 1
    L = O.lib()
    L.orc_phase_times.argtypes = [O._dp]
    R = O.System(d["pos"], d["radius"], d["L"])
    R.first_update()
    L.orc_phase_reset()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        R.step()
    el = time.perf_counter() - t0
    ph = np.zeros(len(PHASES))
    L.orc_phase_times(O.P(ph))
    print(json.dumps({"steps": nsteps, "step_s": el / nsteps,
                      "phases": {k: float(v) / nsteps for k, v in zip(PHASES, ph)}}), flush=True)


if __name__ == "__main__":
    main()
