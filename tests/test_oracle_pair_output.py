"""Per-interaction force columns and the non-affine displacement of the
oracle (the int_ / par_ snapshot quantities, Simulation.cpp:854-1006), and
the snapshot writers (lf_dem_amd/output.py); CPU only.

The reference's own par_/int_ goldens were written by an older version
(v2.0-525) whose velocities and lubrication forces follow other parameter
semantics, so they pin only the geometry: the printed positions (origin at
the box centre, the upper half moved through the Lees-Edwards face), the
pair labels, contact states, normals and gaps.  The force columns are pinned
by what the mechanics fixes:
  * force balance: for every particle, the pair forces (lubrication at the
    non-affine velocities with the R_FE:E term, contact spring + dashpot at
    the total velocities, repulsion) plus the Stokes drag on its non-affine
    velocity sum to zero, because the velocities solve that balance;
  * the contact xF column summed over pairs is the box volume times the
    contact stress's S:E / rate^2.
"""
import numpy as np
import pytest

from lf_dem_amd import output
from oracle import oracle as O
from tests.conftest import conf


def _balance(R, rad):
    po, pr, st = R.pair_output(), R.pairs(), R.state()
    n = pr["nvec"]
    tot = -po[:, [0]]*n + po[:, 1:4] + po[:, [4]]*n + po[:, 5:8] - po[:, [8]]*n
    F = np.zeros((len(rad), 3))
    np.add.at(F, pr["p0"], tot)
    np.add.at(F, pr["p1"], -tot)
    drag = -R.params.sd_coeff*rad[:, None]*st["na_vel"]
    return F+drag, np.abs(drag).max(), po, pr


@pytest.fixture(scope="module")
def c1_stepped():
    pos, rad, L = conf(500, 0.5, 1)
    R = O.System(pos, rad, L)
    R.first_update()
    for _ in range(3):
        R.step()
    R.set_in_predictor(False)
    R.compute_velocities()
    return R, rad, L


def test_pair_forces_balance_every_particle(c1_stepped):
    R, rad, _ = c1_stepped
    res, scale, po, pr = _balance(R, rad)
    assert (pr["state"] > 0).sum() > 100                # contacts (spring + dashpot) are exercised
    assert np.abs(po[:, 0]).max() > 0 and np.abs(po[:, 8]).max() > 0
    assert np.abs(res).max() <= 1e-10*scale


def test_pair_forces_balance_under_stress_control(golden):
    p = O.default_params(kn=5000, kt=2500, sd_coeff=1e-8, tau_n=2e-3, tau_t=0, repulsion=1.0,
                         repulsive_length=0.02, shear_rate=1.0)
    R = O.System(golden["dat_pos"], golden["dat_radius"], golden["dat_L"], shear_disp=golden["dat_disp"], params=p)
    R.set_control("stress", 5.0/(6*np.pi))
    R.first_update()
    R.set_in_predictor(True)
    R.compute_velocities_components()
    res, scale, po, _ = _balance(R, golden["dat_radius"])
    # the drag is 1e-8 of the usual: compare with the lubrication forces instead
    assert np.abs(res).max() <= 1e-10*np.abs(po[:, :4]).max()


def test_a_wrong_lubrication_term_breaks_the_balance(c1_stepped):
    """The balance is sensitive to the force columns: dropping the R_FE:E
    term, or flipping the sign of the tangential part, shows up."""
    R, rad, _ = c1_stepped
    res, scale, po, pr = _balance(R, rad)
    n = pr["nvec"]
    bad = po.copy()
    bad[:, 1:4] *= -1
    tot = -bad[:, [0]]*n + bad[:, 1:4] + bad[:, [4]]*n + bad[:, 5:8] - bad[:, [8]]*n
    F = np.zeros((len(rad), 3))
    np.add.at(F, pr["p0"], tot)
    np.add.at(F, pr["p1"], -tot)
    assert np.abs(F-R.params.sd_coeff*rad[:, None]*R.state()["na_vel"]).max() > 1e-3*scale


def test_contact_xF_column_is_the_contact_stress(c1_stepped):
    R, rad, L = c1_stepped
    R.calc_stress()
    po = R.pair_output()
    S = R.stress()
    rate = R.shear_rate
    E = output.imposed_strain_rate(rate)
    expect = np.prod(L)*output._ddot(S["groups"]["contact"], E)/rate**2
    assert expect != 0
    assert po[:, 9].sum() == pytest.approx(expect, rel=1e-12)


def test_na_disp_accumulates_the_corrector_velocity():
    pos, rad, L = conf(200, 0.45, 3)
    R = O.System(pos, rad, L)
    R.first_update()
    assert np.all(R.na_disp() == 0)
    acc = np.zeros((len(rad), 3))
    for _ in range(3):
        R.step()
        st = R.state()
        acc = acc+st["na_vel"]*st["dt"]
        np.testing.assert_allclose(R.na_disp(), acc, rtol=0, atol=1e-15)


def test_snapshot_geometry_matches_reference_files(golden):
    """t = 0 snapshots of the reference's stress-controlled check: printed
    positions, pair labels, contact states, normals and gaps, to the printed
    digits.  (The rate check's first snapshot was printed from a state that
    neither of the check configurations reproduces, so it is not used.)"""
    pos, rad, L, sd = (golden["dat_%s" % k] for k in ("pos", "radius", "L", "disp"))
    gp, gi = golden["stress_par"], golden["stress_int"]
    shifted = output.shift_up(pos, L, sd)
    np.testing.assert_allclose(shifted, gp[:, 2:5], rtol=0, atol=6e-6*np.abs(gp[:, 2:5]).max())
    assert np.array_equal(gp[:, 0], np.arange(len(rad))) and np.array_equal(gp[:, 1], rad)
    R = O.System(pos, rad, L, shear_disp=sd)
    R.first_update()
    pr = R.pairs()
    rows = output.interaction_rows(pr, np.zeros((len(pr["p0"]), 10)))
    geo = np.array([[r[0][3], r[1][3], r[2][3], *r[3][3], r[4][3]] for r in rows])
    assert np.array_equal(geo[:, :3], gi[:, :3])
    np.testing.assert_allclose(geo[:, 3:7], gi[:, 3:7], rtol=0, atol=5e-6)


def test_shifted_positions_are_lees_edwards_images(golden):
    """With a shear displacement (the rate check's .bin configuration): every
    printed position, moved back by half a box in x and y, is a Lees-Edwards
    image of the particle (the oracle's periodize brings it back into the
    box at the particle's position), and it lies in the centred box."""
    pos, L, sd = golden["bin_pos"], golden["bin_L"], golden["bin_disp"]
    assert sd[0] != 0
    s = output.shift_up(pos, L, sd)
    assert np.all(s >= -0.5*L) and np.all(s <= 0.5*L)
    assert np.any(s[:, 2] < 0) and np.any(s[:, 2] > 0)
    for i in range(len(pos)):
        p = s[i]+np.array([0.5*L[0], 0.5*L[1], 0.0])
        O.lib().orc_periodize(O.P(L), O.P(sd), p.ctypes.data_as(O._dp))
        d = p-pos[i]
        d[:2] -= L[:2]*np.round(d[:2]/L[:2])
        assert np.abs(d).max() < 1e-12, i


def test_snapshot_files_round_trip(tmp_path, c1_stepped):
    """par_ / int_ files: file header, column header once, a snapshot header
    per snapshot, rows read back to the printed digits."""
    R, rad, L = c1_stepped
    st = R.state()
    pr, po = R.pairs(), R.pair_output()
    head = output.data_header(len(rad), 0.5, L)
    fp = output.OutputFile(str(tmp_path/"par_x.dat"), head, "hydro")
    fi = output.OutputFile(str(tmp_path/"int_x.dat"), head, "hydro")
    for snap in range(2):
        h = output.snapshot_header(st["strain"], st["shear_disp"][0], R.shear_rate)
        fp.write_snapshot(h, output.particle_rows(st, rad, L, R.na_disp(), R.shear_rate))
        fi.write_snapshot(h, output.interaction_rows(pr, po))
    fi.write_snapshot(h, [])
    fp.close()
    fi.close()
    text = open(tmp_path/"par_x.dat").read()
    assert text.count("# data in hydro units.") == 1 and text.count("# cumulated strain : ") == 2
    assert "#3-5: position (x, y, z)" in text and "#12-14: non affine displacement (x, y, z)" in text
    names, rows = output.read_rows(str(tmp_path/"int_x.dat"))
    assert len(rows) == 2*len(pr["p0"]) and len(names) == 17
    np.testing.assert_allclose(rows[:len(pr["p0"]), 7], po[:, 0], rtol=1e-5, atol=1e-300)
    np.testing.assert_allclose(rows[:len(pr["p0"]), 11], np.abs(po[:, 4]), rtol=1e-5)
