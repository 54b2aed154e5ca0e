"""LF_DEM's data_ and st_ output files (Simulation::outputData,
Simulation.cpp:641-768; OutputData.h:27-207; header of
Simulation::createDataHeader, Simulation.cpp:812-824) and the par_ / int_
snapshot files (Simulation::outputParFileTxt / outputIntFileTxt,
Simulation.cpp:854-1006; snapshot header of getSnapshotHeader,
Simulation.cpp:770-794).

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


def vel_difference(shear_rate, Lz, theta_shear=0.0):
    """System::setVelocityDifference (System.cpp:1967-1970): 2 E_inf . (0, 0, Lz)."""
    return shear_rate * Lz * np.array([math.cos(theta_shear), math.sin(theta_shear), 0.0])


def shift_up(pos, L, shear_disp):
    """Positions as the snapshot files print them with origin_zero_flow
    (Simulation.cpp:866-872, shiftUpCoordinate :795-810): the origin at the
    box centre, the particles of the upper half moved one box down through
    the Lees-Edwards face (x and y shifted back by the shear displacement,
    wrapped once)."""
    L = np.asarray(L, dtype=float)
    x = np.array(pos, dtype=float) - 0.5 * L
    z = x[:, 2] + 0.5 * L[2]
    up = z > 0.5 * L[2]
    x[up, 0] -= shear_disp[0]
    x[up, 1] -= shear_disp[1]
    x[up & (x[:, 0] < -0.5 * L[0]), 0] += L[0]
    x[up & (x[:, 1] < -0.5 * L[1]), 1] += L[1]
    z[up] -= L[2]
    x[:, 2] = z
    return x


def particle_rows(state, radius, L, na_disp, shear_rate, theta_shear=0.0, origin_zero_flow=True):
    """The rows of one par_ snapshot (Simulation::outputParFileTxt,
    Simulation.cpp:854-939; simple shear, no optional columns): index,
    radius, position, velocity (minus the Lees-Edwards velocity difference
    below the shifted origin), angular velocity, non-affine displacement.
    state: System.state(), na_disp: System.na_disp(), shear_rate: the
    current one (System.stress()["shear_rate"] or the imposed rate)."""
    pos = np.asarray(state["pos"], dtype=float)
    vel = np.array(state["vel"], dtype=float)
    if origin_zero_flow:
        pos = shift_up(pos, L, state["shear_disp"])
        vel[pos[:, 2] < 0] -= vel_difference(shear_rate, L[2], theta_shear)
    ang = np.asarray(state["ang_vel"], dtype=float)
    rows = []
    for i in range(len(radius)):
        rows.append([("particle index", "none", 1, i),
                     ("radius", "none", 1, radius[i]),
                     ("position (x, y, z)", "none", 3, tuple(pos[i])),
                     ("velocity (x, y, z)", "velocity", 3, tuple(vel[i])),
                     ("angular velocity (x, y, z)", "none", 3, tuple(ang[i])),
                     ("non affine displacement (x, y, z)", "none", 3, tuple(na_disp[i]))])
    return rows


CONTACT_STATE = ("contact state (0 = no contact, 1 = frictionless contact, 2 = non-sliding frictional, "
                 "3 = sliding frictional)")


def interaction_rows(pairs, pair_output):
    """The rows of one int_ snapshot (Simulation::outputIntFileTxt,
    Simulation.cpp:941-1006, with lubrication): pairs from System.pairs(),
    pair_output from System.pair_output() (same order)."""
    rows = []
    o = np.asarray(pair_output, dtype=float)
    for k in range(len(pairs["p0"])):
        rows.append([("particle 1 label", "none", 1, int(pairs["p0"][k])),
                     ("particle 2 label", "none", 1, int(pairs["p1"][k])),
                     (CONTACT_STATE, "none", 1, int(pairs["state"][k])),
                     ("normal vector, oriented from particle 1 to particle 2", "none", 3, tuple(pairs["nvec"][k])),
                     ("dimensionless gap = s-2, s = 2r/(a1+a2)", "none", 1, pairs["gap"][k]),
                     ("normal part of the lubrication force (positive for compression)", "force", 1, o[k, 0]),
                     ("tangential part of the lubrication force", "force", 3, tuple(o[k, 1:4])),
                     ("norm of the normal part of the contact force", "force", 1, abs(o[k, 4])),
                     ("tangential part of the contact force", "force", 3, tuple(o[k, 5:8])),
                     ("norm of the normal repulsive force", "force", 1, o[k, 8]),
                     ("Viscosity contribution of contact xF", "stress", 1, o[k, 9])])
    return rows


def snapshot_header(strain, shear_disp_x, rate, target_stress=None):
    """Simulation::getSnapshotHeader (Simulation.cpp:770-794), simple shear:
    rate is Simulation::getRate (the input rate under rate control, the
    current shear rate under stress control); target_stress the input value
    under stress control."""
    s = "# cumulated strain : %s\n# shear disp : %s\n# shear rate : %s\n" % (fmt(strain), fmt(shear_disp_x),
                                                                           fmt(rate))
    if target_stress is not None:
        s += "# target stress : %s\n" % fmt(target_stress)
    return s


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
    one line per row; a snapshot file puts a snapshot header before each
    snapshot's rows."""

    def __init__(self, path, header, unit, ratio=1.0):
        self.f = open(path, "w")
        self.f.write(header)
        self.unit = unit
        self.ratio = ratio          # internal force unit / output force unit
        self.first = True

    def _column_header(self, cols):
        self.f.write("# data in %s units.\n" % self.unit)
        i = 1
        for name, _, width, _ in cols:
            self.f.write(("#%d: %s\n" % (i, name)) if width == 1 else ("#%d-%d: %s\n" % (i, i + width - 1, name)))
            i += width
        self.f.write("\n")
        self.first = False

    def _line(self, cols):
        out = []
        for _, dim, width, v in cols:
            c = _CONVERT[dim](self.ratio)
            vals = [v] if width == 1 else list(v)
            out += [fmt(c * x) for x in vals]
        return " ".join(out) + " \n"

    def write(self, cols):
        if self.first:
            self._column_header(cols)
        self.f.write(self._line(cols))
        self.f.flush()

    def write_snapshot(self, header, rows):
        """OutputData::writeToFile(header) of a snapshot file: the column
        header before the first snapshot, then the snapshot header and one
        line per row.  Nothing is written for an empty snapshot (the
        interaction file without interactions, Simulation.cpp:1002-1004)."""
        if not rows:
            return
        if self.first:
            self._column_header(rows[0])
        self.f.write(header)
        for cols in rows:
            self.f.write(self._line(cols))
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
