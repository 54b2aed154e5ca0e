"""The oracle's step loop against the reference's multi-step outputs (CPU).

Pins orc_time_evolution / orc_step_ex (System::timeEvolution,
timeEvolutionPredictorCorrectorMethod, adaptTimeStep, timeStepBoxing:
System.cpp:888-983, 1013-1066, 1124-1187, 1203-1279, 672-717) and, through
them, the pair lifecycle over steps, stress control and the fixed_dt path:
every row of LF_DEM's own data_ files for its two regression runs (11 rows
each, 1e-3 strain apart) is reproduced to the printed 6 digits in every
column that the current Simulation::outputData shares with the file:
time, strain, shear rate, viscosity and its groups, N1, N2, particle
pressure, energy, min gap, contact numbers, number of interactions (745 ->
746 -> 745 in the stress run, 745 -> 746 -> 747 in the rate run), max
velocity, max angular velocity, average dt and the Lees-Edwards shear
displacement.  The row dt column (avg_dt) pins the adaptive step sequence:
a wrong disp_max rule, midpoint average or end-strain clamp moves it.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import golden_rows as G


def oracle_system(tag):
    c, S = G.setup(tag)
    R = O.System(c.pos, c.radius, c.L, shear_disp=c.shear_disp, params=O.default_params(**S.params))
    if S.control == "stress":
        R.set_control("stress", S.target_stress)
    R.first_update()
    return R, c, S


def test_stress_controlled_run_reproduces_all_golden_rows():
    R, c, S = oracle_system("stress5r")
    _, g = G.golden("stress5r")
    produced = G.rows(R, c, S, g[1:, 1])          # rows 1..10 at strain 0.001 .. 0.01
    assert [p[2] for p in produced] == [4, 4, 4, 4, 4, 4, 3, 3, 3, 3]
    assert G.mismatches("stress5r", produced, 1) == []
    assert [p[0]["number of interaction"] for p in produced][5:8] == [745, 746, 745]


def test_fixed_dt_rate_run_reproduces_all_golden_rows():
    R, c, S = oracle_system("rate02r")
    assert S.fixed_dt and S.output_ratio == pytest.approx(0.2, rel=1e-15)
    _, g = G.golden("rate02r")
    produced = G.rows(R, c, S, g[:, 1])           # rows 0..10 at strain 5e-5, 1.05e-3, ...
    assert [p[2] for p in produced] == [1] + [20] * 10
    assert G.mismatches("rate02r", produced, 0) == []


def test_planted_step_rule_mistakes_are_caught():
    """The row comparison fails when the dt rule or the final-step
    convention is changed in the oracle: driving the stress run with a
    larger displacement cap, or to an end strain one step short."""
    _, g = G.golden("stress5r")
    for kw, ends in (({"disp_max": 1.1e-3}, g[1:4, 1]), ({}, g[1:4, 1] - 2e-4)):
        c, S = G.setup("stress5r")
        p = dict(S.params, **kw)
        R = O.System(c.pos, c.radius, c.L, shear_disp=c.shear_disp, params=O.default_params(**p))
        R.set_control("stress", S.target_stress)
        R.first_update()
        assert G.mismatches("stress5r", G.rows(R, c, S, ends), 1) != []


def test_data_writer_rows_read_back_as_the_reference_prints_them(tmp_path):
    """lf_dem_amd.output writes data_ / st_ files in OutputData's layout
    (file header, `# data in <unit> units.`, `#k: name` / `#k-m: name`
    column lines, a blank line, rows of 6-digit values); the stress run's
    rows read back from the written file match the reference's rows."""
    from lf_dem_amd import output
    R, c, S = oracle_system("stress5r")
    _, g = G.golden("stress5r")
    path, spath = str(tmp_path / "data_x.dat"), str(tmp_path / "st_x.dat")
    hdr = output.data_header(200, c.vf, c.L)
    f = output.OutputFile(path, hdr, S.output_unit, S.output_ratio)
    fs = output.OutputFile(spath, hdr, S.output_unit, S.output_ratio)
    for s_end in g[1:4, 1]:
        t0 = R.state()["time"]
        n = R.time_evolution(s_end)
        st, sr = R.state(), R.stress()
        f.write(output.data_columns(sr, st, R.pairs(), c.radius, S.params,
                                    output.average_dt(t0, st["time"], st["dt"], n)))
        fs.write(output.stress_columns(sr, st))
    f.close()
    fs.close()
    text = open(path).read().splitlines()
    assert text[:8] == ["# LF_DEM version lf_dem_amd", "# np 200", "# VF 0.45", "# Lx 13.9626", "# Ly 13.9626",
                        "# Lz 13.9626", "# flow_type shear", "# data in repulsion units."]
    assert text[8:12] == ["#1: time", "#2: cumulated shear strain", "#3: shear rate", "#4: viscosity"]
    assert text[-4] == "" and text[-5] == "#28-30: shear strain"
    names, rows = output.read_rows(path)
    col = {n: i for i, n in enumerate(names)}
    gnames, _ = G.golden("stress5r")
    gcol = {n: i for i, n in enumerate(gnames)}
    for k in range(3):
        for ours, theirs in G.SAME.items():
            assert G.printed(rows[k, col[ours]], g[1 + k, gcol[theirs]]), (k, ours)
        assert rows[k, col["shear strain[0]"]] == pytest.approx(g[1 + k, 1], rel=1e-12)
    snames, srows = output.read_rows(spath)
    assert snames[:4] == ["time", "cumulated strain", "shear rate", "total stress tensor (xx, xy, xz, yz, yy, zz)[0]"]
    assert srows.shape == (3, 3 + 6 * 5)
    # xz component of the total stress over the shear rate, times 6 pi: the viscosity column (times 2 E_xz = rate)
    assert srows[:, 5] / srows[:, 2] == pytest.approx(rows[:, col["viscosity"]], rel=2e-5)
