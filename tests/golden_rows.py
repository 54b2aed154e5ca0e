"""Shared driver for the reference's multi-step regression outputs
(tests/golden/data_rows.npz, written by tests/golden/make_data_rows.py from
LF_DEM/check/reference/data_*_repulsive_{stress_stress5r,rate0.2r}.dat).

Each golden row k is the data line LF_DEM printed after one call of
System::timeEvolution (System.cpp:1203-1279) ending at the row's strain:
steps while keepRunning, then the step that computes the stress.  The
driver calls time_evolution(strain_k) on a system (the oracle's or the HIP
library's), forms the row with lf_dem_amd.output.data_columns and compares
every column that LF_DEM's current Simulation::outputData shares with the
file's (older) layout, to the 6 printed digits.

The files come from LF_DEM v2.0-636.  Two of their conventions differ from
the current source, and the driver follows the current source:
  * fixed_dt runs (the rate check): the older loop had no
    `loop_time_adjust` (System.cpp:1237-1242) and so stopped one step later;
    the output clock at time 1e-3 k then printed the row at strain
    1e-3 k + dt.  Calling the current loop with the row's own strain as the
    end makes the same steps (the row strains are the targets).
  * "max velocity" and "max angular velocity" of the fixed_dt (rate) file:
    the current source never updates System::max_velocity with fixed_dt
    (System.cpp:2309), and both columns of that file sit 4e-5 to 1e-4 relative
    from the current definitions (maximum non-affine velocity and
    evaluateMaxAngVelocity of the last solve) in all 11 rows, with opposite
    signs, while every stress, pair and geometry column agrees to 6 digits:
    the older code took them at another stage of the step.  There the
    largest non-affine velocity of the last solve and the max angular
    velocity are compared to 2e-4 relative; the stress file (no fixed_dt)
    pins both columns to its 6 digits.
"""
import math
import os

import numpy as np

from lf_dem_amd import formats, output, params

HERE = os.path.dirname(os.path.abspath(__file__))
CHECK = os.path.join(HERE, "golden", "lfdem_check")
STEM = os.path.join(CHECK, "D3N200VF0.45Bidi1.4_0.5Cubic_1")

# current outputData name -> the golden file's name for the same quantity
SAME = {"time": "time", "cumulated shear strain": "shear strain", "shear rate": "shear rate",
        "viscosity": "viscosity", "Viscosity(hydro)": "Viscosity(lub)", "N1 viscosity": "N1 viscosity",
        "N2 viscosity": "N2 viscosity", "particle pressure": "particle pressure",
        "particle pressure contact": "particle pressure contact", "energy": "energy", "min gap": "min gap",
        "max tangential displacement": "max tangential displacement",
        "max rolling displacement": "max rolling displacement", "contact number": "contact number",
        "frictional contact number": "frictional contact number",
        "number of interaction": "number of interaction", "max velocity": "max velocity",
        "max angular velocity": "max angular velocity", "dt": "dt", "kn": "kn", "kt": "kt", "kr": "kr"}
# current groups that are the sum of two older columns (GU and xF of the same force)
SUMS = {"Viscosity(contact)": ("Viscosity(xF_contact part)", "Viscosity(GU_contact part)"),
        "Viscosity(repulsion)": ("Viscosity(repulsive force XF)", "Viscosity(repulsive force GU)")}

RUNS = {"stress5r": ("dat", "repulsive_stress.txt", "stress", "5r"),
        "rate02r": ("bin", "repulsive.txt", "rate", "0.2r")}


def golden(tag):
    d = np.load(os.path.join(HERE, "golden", "data_rows.npz"))
    return [str(n) for n in d[tag + "_names"]], d[tag + "_rows"]


def setup(tag):
    """(configuration, resolved parameters) of the reference's check run."""
    ext, pfile, control, value = RUNS[tag]
    c = formats.read(STEM + "." + ext)
    S = params.load(os.path.join(CHECK, pfile), control, value, legacy_keys=True)
    return c, S


def rows(system, c, S, strains):
    """Drives `system` (oracle.System or capi.System, created on c with
    S.params, control set, first update done) through time_evolution to each
    strain; returns per row the dict name -> value in output units, the
    shear displacement and the number of steps."""
    out = []
    for s_end in strains:
        t0 = system.state()["time"]
        n = system.time_evolution(s_end)
        st = system.state()
        avg = output.average_dt(t0, st["time"], st["dt"], n)
        cols = output.data_columns(system.stress(), st, system.pairs(), c.radius, S.params, avg)
        row = {}
        for name, dim, width, v in cols:
            if width == 1:
                row[name] = output._CONVERT[dim](S.output_ratio) * v
        # the largest non-affine velocity of the last solve, in output units
        row["max |na velocity|"] = S.output_ratio * float(np.max(np.linalg.norm(st["na_vel"], axis=1)))
        out.append((row, st["shear_disp"].copy(), n))
    return out


def printed(x, g):
    """|x - g| within half a unit of the 6th significant digit of g."""
    if g == 0:                      # printed "0" or "-0": exactly zero (setprecision prints 1e-20 as such)
        return x == 0
    ulp6 = 10.0 ** (math.floor(math.log10(abs(g))) - 5)
    return abs(x - g) <= 0.5 * ulp6 * (1 + 1e-9)


def mismatches(tag, produced, first_row):
    """Columns of the produced rows that differ from the golden rows
    first_row, first_row+1, ... beyond the printed digits."""
    names, g = golden(tag)
    col = {n: i for i, n in enumerate(names)}
    bad = []
    for k, (row, sd, _) in enumerate(produced):
        gr = g[first_row + k]
        for ours, theirs in SAME.items():
            if tag == "rate02r" and ours in ("max velocity", "max angular velocity"):
                mine = row["max |na velocity|" if ours == "max velocity" else ours]
                if abs(mine - gr[col[theirs]]) > 2e-4 * abs(gr[col[theirs]]):
                    bad.append((first_row + k, ours, mine, gr[col[theirs]]))
                continue
            if not printed(row[ours], gr[col[theirs]]):
                bad.append((first_row + k, ours, row[ours], gr[col[theirs]]))
        for ours, (a, b) in SUMS.items():
            ref = gr[col[a]] + gr[col[b]]
            tol = 0.5 * sum(10.0 ** (math.floor(math.log10(abs(v))) - 5) for v in (gr[col[a]], gr[col[b]]) if v)
            if abs(row[ours] - ref) > tol * (1 + 1e-9) + 1e-300:
                bad.append((first_row + k, ours, row[ours], ref))
        for j, name in enumerate(("shear displacement x", "shear displacement y")):
            if not printed(sd[j], gr[col[name]]):
                bad.append((first_row + k, name, sd[j], gr[col[name]]))
    return bad
