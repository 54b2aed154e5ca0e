"""LF_DEM's data_ and st_ output files (Simulation::outputData,
Simulation.cpp:641-768; OutputData.h:27-207; header of
Simulation::createDataHeader, Simulation.cpp:812-824).

A row is formed from plain values: the stress dict of System.stress() (group
and total stress tensors in the internal unit, shear rate, viscosity), the
state dict of System.state(), the pair dict of System.pairs(), the radii, the
parameters and the average step of the last time evolution.  Values are
converted to the output unit as OutputData::setDimensionlessNumber does
(OutputData.h:118-127): viscosities times 6 pi, stresses times 6 pi times
the internal / output force ratio, times divided by it, rates and velocities
multiplied by it.  Numbers are printed with 6 significant digits
(std::setprecision(6)), each followed by a space.

Columns of this library's scope (no walls, no Brownian force, no cohesion,
simple shear): the entries of Simulation.cpp:677-740 in their order.
"""
import math

import numpy as np

from . import observables as ob

GROUPS = ("contact", "dashpot", "hydro", "repulsion")   # std::map order of total_stress_groups

# OutputData::setDimensionlessNumber (OutputData.h:118-127)
_CONVERT = {"none": lambda r: 1.0, "viscosity": lambda r: 6 * math.pi, "stress": lambda r: 6 * math.pi * r,
            "force": lambda r: r, "time": lambda r: 1.0 / r, "rate": lambda r: r, "velocity": lambda r: r}


def _ddot(a, b):
    """Sym2Tensor doubledot of (xx, xy, xz, yz, yy, zz) tensors."""
    return a[0]*b[0] + 2*a[1]*b[1] + 2*a[2]*b[2] + 2*a[3]*b[3] + a[4]*b[4] + a[5]*b[5]


def imposed_strain_rate(shear_rate, theta_shear=0.0):
    """E_infinity of simple shear (System::setShearDirection, System.cpp:2006-2015)."""
    c, s = math.cos(theta_shear), math.sin(theta_shear)
    return shear_rate * np.array([0.0, 0.0, c / 2, s / 2, 0.0, 0.0])


def average_dt(t0, t1, dt_last, nsteps):
    """System::avg_dt after timeEvolution (System.cpp:1256-1265): the mean
    step of the loop, which excludes the final stress step; the current dt
    when the loop made no step.  t0 / t1 are the times before / after the
    evolution, dt_last the step of its final stress step, nsteps the steps it
    made in total."""
    if nsteps <= 1:
        return dt_last
    return (t1 - dt_last - t0) / (nsteps - 1)


def data_columns(stress, state, pairs, radius, params, avg_dt):
    """(name, physical dimension, width, value) of each entry of one data_ row."""
    sr = stress["shear_rate"]
    theta = params.get("theta_shear", 0.0)
    total = np.asarray(stress["total"], dtype=float)
    groups = {g: np.asarray(stress["groups"][g], dtype=float) for g in GROUPS}
    if sr != 0:
        E = imposed_strain_rate(sr, theta)
        s0 = np.array([-0.25, 0, 0, 0, 0.5, -0.25])       # stress_basis_0 (SimulationInit.cpp:659-661)
        s3 = np.array([-0.5, 0, 0, 0, 0, 0.5])            # stress_basis_3
        visc = 0.5 * _ddot(total, E) / _ddot(E, E)
        f0 = _ddot(total, s0) / _ddot(s0, s0)
        f3 = _ddot(total, s3) / _ddot(s3, s3)
    else:
        E = imposed_strain_rate(1.0, theta)
        visc = 0.5 * _ddot(total, E) / _ddot(E, E)
        f0 = f3 = 0.0
    cols = [("time", "time", 1, state["time"]),
            ("cumulated shear strain", "none", 1, state["strain"]),
            ("shear rate", "rate", 1, sr),
            ("viscosity", "viscosity", 1, visc)]
    for g in GROUPS:
        cols.append(("Viscosity(%s)" % g, "viscosity", 1, 0.5 * _ddot(groups[g], E) / _ddot(E, E)))
    cols += [("inviscid function 0th", "viscosity", 1, f0),
             ("inviscid function 3rd", "viscosity", 1, f3),
             ("N1 viscosity", "viscosity", 1, (total[0] - total[5]) / sr if sr else float("inf")),
             ("N2 viscosity", "viscosity", 1, (total[5] - total[4]) / sr if sr else float("inf")),
             ("particle pressure", "stress", 1, -(total[0] + total[4] + total[5]) / 3),
             ("particle pressure contact", "stress", 1,
              -(groups["contact"][0] + groups["contact"][4] + groups["contact"][5]) / 3)]
    np_ = len(radius)
    energy = ob.potential_energy(pairs, radius, params["kn"], params["kt"], params["repulsion"],
                                 params["repulsive_length"])
    cn, fcn = ob.contact_numbers(pairs, np_)
    ang = np.asarray(state["na_ang_vel"])
    cols += [("energy", "none", 1, energy),
             ("min gap", "none", 1, ob.min_gap(pairs, radius, 2 + params["lub_max_gap"])),
             ("max tangential displacement", "none", 1, ob.max_disp_tan(pairs)),
             ("max rolling displacement", "none", 1, 0.0),
             ("contact number", "none", 1, cn),
             ("frictional contact number", "none", 1, fcn),
             ("number of interaction", "none", 1, len(pairs["p0"])),
             ("max velocity", "velocity", 1, state["max_velocity"]),
             ("max angular velocity", "velocity", 1,
              float(np.max(np.linalg.norm(ang, axis=1))) if np_ else 0.0),   # evaluateMaxAngVelocity
             ("dt", "time", 1, avg_dt),
             ("kn", "none", 1, params["kn"]),
             ("kt", "none", 1, params["kt"]),
             ("kr", "none", 1, 0.0)]
    strain = state["strain"]
    cols.append(("shear strain", "none", 3, (strain * math.cos(theta), strain * math.sin(theta), 0.0)))
    return cols


def stress_columns(stress, state):
    """Entries of one st_ row (Simulation.cpp:757-767)."""
    cols = [("time", "time", 1, state["time"]),
            ("cumulated strain", "none", 1, state["strain"]),
            ("shear rate", "rate", 1, stress["shear_rate"]),
            ("total stress tensor (xx, xy, xz, yz, yy, zz)", "stress", 6, tuple(stress["total"]))]
    for g in GROUPS:
        cols.append(("%s stress tensor (xx, xy, xz, yz, yy, zz)" % g, "stress", 6, tuple(stress["groups"][g])))
    return cols


def fmt(x):
    """A double as std::ostream prints it with setprecision(6)."""
    x = float(x)
    if math.isnan(x):
        return "-nan" if math.copysign(1, x) < 0 else "nan"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    return "%.6g" % x


def data_header(np_, vf, L, flow_type="shear", version="lf_dem_amd"):
    """Simulation::createDataHeader (Simulation.cpp:812-824)."""
    return ("# LF_DEM version %s\n# np %d\n# VF %s\n# Lx %s\n# Ly %s\n# Lz %s\n# flow_type %s\n"
            % (version, np_, fmt(vf), fmt(L[0]), fmt(L[1]), fmt(L[2]), flow_type))


class OutputFile:
    """OutputData: the file header once, the column header on the first
    row (`# data in <unit> units.`, then `#k: name` or `#k-m: name`), then
    one line per row."""

    def __init__(self, path, header, unit, ratio=1.0):
        self.f = open(path, "w")
        self.f.write(header)
        self.unit = unit
        self.ratio = ratio          # internal force unit / output force unit
        self.first = True

    def write(self, cols):
        if self.first:
            self.f.write("# data in %s units.\n" % self.unit)
            i = 1
            for name, _, width, _ in cols:
                self.f.write(("#%d: %s\n" % (i, name)) if width == 1 else ("#%d-%d: %s\n" % (i, i + width - 1, name)))
                i += width
            self.f.write("\n")
            self.first = False
        out = []
        for _, dim, width, v in cols:
            c = _CONVERT[dim](self.ratio)
            vals = [v] if width == 1 else list(v)
            out += [fmt(c * x) for x in vals]
        self.f.write(" ".join(out) + " \n")
        self.f.flush()

    def close(self):
        self.f.close()


def read_rows(path):
    """Column names (expanded per width: name, name[1], ...) and rows of a
    data_ / st_ file written by LF_DEM or by OutputFile."""
    names, rows = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                head = line[1:].split(":", 1)
                idx = head[0].strip()
                if len(head) == 2 and idx.replace("-", "").isdigit():
                    lo, _, hi = idx.partition("-")
                    w = int(hi) - int(lo) + 1 if hi else 1
                    name = head[1].strip()
                    names += [name] if w == 1 else ["%s[%d]" % (name, k) for k in range(w)]
                continue
            if line.strip():
                rows.append([float(x) for x in line.split()])
    return names, np.array(rows)
