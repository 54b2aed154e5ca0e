"""Writes tests/golden/rheology_n200.npz: the oracle's viscosity against strain
over one strain unit for the reference's N=200 configuration
(check/confs/D3N200VF0.45Bidi1.4_0.5Cubic_1.dat, as stored in check_n200.npz),
rate-controlled shear with the default parameters, outputs every 0.05 strain.

The oracle is deterministic; a 1e-9 perturbation of one coordinate changes
these viscosities by at most 1e-10 relative over the whole unit of strain, so
the series is a reproducible reference for the GPU path.  About a minute of
CPU time.

usage: python tests/golden/make_rheology.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

STRAIN_STEP = 0.05
OUTPUTS = 20


def series(pos, rad, L, sd, outputs=OUTPUTS):
    R = O.System(pos, rad, L, sd)
    R.first_update()
    visc, steps, strain = [], [], []
    for k in range(1, outputs + 1):
        steps.append(R.time_evolution(STRAIN_STEP * k))
        visc.append(R.viscosity)
        strain.append(R.state()["strain"])
    return np.array(visc), np.array(steps), np.array(strain)


def main():
    g = np.load(os.path.join(HERE, "check_n200.npz"))
    visc, steps, strain = series(g["dat_pos"], g["dat_radius"], g["dat_L"], g["dat_disp"])
    np.savez(os.path.join(HERE, "rheology_n200.npz"), viscosity=visc, steps=steps, strain=strain,
             strain_step=STRAIN_STEP)
    print("viscosity", visc)
    print("steps", steps.sum())


if __name__ == "__main__":
    main()
