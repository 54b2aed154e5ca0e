"""Pins of the oracle's contact history and dashpot right-hand side by what
the mechanics of rigid spheres fixes (CPU).

* Tangential stretch (Contact::incrementTangentialDisplacement,
  Contact.cpp:125-160, restated in contact_increment): over a step dt the
  stretch grows by dt times the relative velocity of the two contact points,
  projected on the tangent plane.  The contact points' velocities are formed
  here from rigid-body kinematics, v + w x (point - centre), with the far
  particle's affine velocity evaluated at its unwrapped position across the
  Lees-Edwards boundary (no z_offset bookkeeping), and the predictor /
  corrector convention: the corrector starts again from the stretch the
  predictor started from.
* Dashpot right-hand side (System::setDashpotForceToParticle,
  ContactDashpot.cpp:426-441, restated in the dashpot component of the
  right-hand side): the force on each particle is minus the dense 12x12
  dashpot resistance of every contact pair times the imposed flow (U_inf at
  both centres, the far one unwrapped, Omega_inf for both rotations),
  summed per particle.  The torques are the reference's own convention,
  a_i n x F_0 on both particles (getRFU_Uinf, ContactDashpot.cpp:426-441):
  the moment of the pair force about each centre, not the torque rows of
  the dense resistance (which add the rotation-rotation C blocks); the test
  states both, so that a change of convention in either direction fails.
"""
import numpy as np
import pytest

from tests.conftest import conf

E_RATE1 = np.array([[0, 0, 0.5], [0, 0, 0], [0.5, 0, 0]])     # simple shear, unit rate (System.cpp:2006-2015)
OMEGA1 = np.array([0, 0.5, 0])


def u_inf(x, rate=1.0):
    return rate * (E_RATE1 @ x + np.cross(OMEGA1, x))


def wrap(oracle, L, sd, x):
    """Position of an unwrapped point inside the Lees-Edwards box and the
    number of z-periods it crossed."""
    x = np.array(x, dtype=float)
    z = oracle.lib().orc_periodize(oracle.P(L), oracle.P(sd), x.ctypes.data_as(oracle._dp))
    return x, z


def pair_12(oracle, db0, db1, od):
    R = np.zeros((12, 12))
    R[:6, :6] = oracle.dblock_dense(db0)
    R[6:, 6:] = oracle.dblock_dense(db1)
    M = oracle.odblock_dense(od)
    R[6:, :6] = M
    R[:6, 6:] = M.T
    return R


def contact_pair(oracle, mu=1e9):
    """Two spheres (a = 1.0 and 1.4) in contact across the top z face of a
    sheared Lees-Edwards box: particle 1's image sits above the box."""
    L = np.array([20.0, 20.0, 20.0])
    sd = np.array([3.3, 0.0, 0.0])
    a = np.array([1.0, 1.4])
    n = np.array([0.3, 0.2, 0.93])
    n /= np.linalg.norm(n)
    x0 = np.array([5.0, 5.0, 19.3])
    x1u = x0 + 2.385 * n                         # overlap 0.015
    x1, zs = wrap(oracle, L, sd, x1u)
    assert zs != 0 and x1[2] < 2                 # the image crossed the top face
    p = oracle.default_params(shear_rate=1.0, mu_static=mu, mu_dynamic=mu, sd_coeff=1.0)
    S = oracle.System(np.array([x0, x1]), a, L, shear_disp=sd, params=p)
    S.first_update()
    pr = S.pairs()
    assert len(pr["p0"]) == 1 and pr["state"][0] >= 2
    assert np.allclose(pr["nvec"][0], n, rtol=0, atol=1e-13)
    return S, a, x0, x1, x1u, n, p


def surface_slip(a, x0, x1, x1u, n, na, w):
    """Relative velocity of the contact points, tangential part: particle 0
    at x0, particle 1's image at x1u (its stored position x1 is wrapped),
    non-affine velocities na (2 x 3), angular velocities w (2 x 3)."""
    v0 = u_inf(x0) + na[0]
    v1u = u_inf(x1u) + na[1]
    c0 = v0 + np.cross(w[0], a[0] * n)
    c1 = v1u + np.cross(w[1], -a[1] * n)
    d = c1 - c0
    return d - np.dot(d, n) * n


def set_total(S, x0, x1, na, w):
    """The oracle's velocity arrays hold total velocities at the stored
    (wrapped) positions."""
    vel = np.array([u_inf(x0) + na[0], u_inf(x1) + na[1]])
    S.set_velocities(vel, np.array(w))


def test_tangential_stretch_is_the_contact_points_slip_across_lees_edwards(oracle):
    S, a, x0, x1, x1u, n, p = contact_pair(oracle)
    rng = np.random.default_rng(7)
    dt = 1e-3
    S.set_dt(dt)
    # predictor: from zero stretch
    na1, w1 = rng.normal(size=(2, 3)), rng.normal(size=(2, 3))
    set_total(S, x0, x1, na1, w1)
    S.set_in_predictor(True)
    S.update_interactions()
    xi1 = dt * surface_slip(a, x0, x1, x1u, n, na1, w1)
    got = S.pairs()["disp_tan"][0]
    assert np.allclose(got, xi1, rtol=0, atol=1e-13 * np.abs(xi1).max())
    # corrector: again from the stretch the predictor started from (zero)
    na2, w2 = rng.normal(size=(2, 3)), rng.normal(size=(2, 3))
    set_total(S, x0, x1, na2, w2)
    S.set_in_predictor(False)
    S.update_interactions()
    xi2 = dt * surface_slip(a, x0, x1, x1u, n, na2, w2)
    got = S.pairs()["disp_tan"][0]
    assert np.allclose(got, xi2, rtol=0, atol=1e-13 * np.abs(xi2).max())
    # next predictor: from the corrector's stretch
    na3, w3 = rng.normal(size=(2, 3)), rng.normal(size=(2, 3))
    set_total(S, x0, x1, na3, w3)
    S.set_in_predictor(True)
    S.update_interactions()
    xi3 = xi2 + dt * surface_slip(a, x0, x1, x1u, n, na3, w3)
    got = S.pairs()["disp_tan"][0]
    assert np.allclose(got, xi3, rtol=0, atol=1e-13 * np.abs(xi3).max())
    # the affine velocity jump across the face is what makes the slip large:
    # without it (far particle taken at its wrapped position) the result differs
    wrong = xi2 + dt * surface_slip(a, x0, x1, x1, n, na3, w3)
    assert np.abs(wrong - got).max() > 1e3 * np.abs(xi3 - got).max() + 1e-6


def test_sliding_contact_keeps_the_slip_direction_at_the_coulomb_limit(oracle):
    """With dt > 0 and a small friction coefficient the stretch is rescaled
    along its own direction to |kt' xi| = mu |Fn| (Contact.cpp:255-291)."""
    mu = 0.05
    S, a, x0, x1, x1u, n, p = contact_pair(oracle, mu=mu)
    rng = np.random.default_rng(3)
    dt = 1e-3
    S.set_dt(dt)
    na, w = 50 * rng.normal(size=(2, 3)), rng.normal(size=(2, 3))
    set_total(S, x0, x1, na, w)
    S.set_in_predictor(True)
    S.update_interactions()
    pr = S.pairs()
    xi = pr["disp_tan"][0]
    slip = surface_slip(a, x0, x1, x1u, n, na, w)
    ro2 = (a[0] + a[1]) / 2
    kn_s, kt_s = ro2 * ro2 * p.kn, ro2 * p.kt          # Contact::setSpringConstants (Contact.cpp:35-45)
    fn = -kn_s * pr["gap"][0]
    assert pr["state"][0] == 3
    assert kt_s * np.linalg.norm(xi) == pytest.approx(mu * fn, rel=1e-12)
    assert np.allclose(xi / np.linalg.norm(xi), slip / np.linalg.norm(slip), rtol=0, atol=1e-12)


def test_dashpot_rhs_is_minus_the_dashpot_resistance_times_the_imposed_flow(oracle):
    pos, rad, L = conf(300, 0.56, 11)
    S = oracle.System(pos, rad, L, shear_disp=(2.7, 0.0, 0.0))
    S.first_update()
    for _ in range(2):
        S.step()
    pr = S.pairs()
    st = S.state()
    contact = np.flatnonzero(pr["state"] >= 1)
    assert len(contact) > 20
    p = S.params
    x = st["pos"]
    F = np.zeros((S.np, 6))
    crossing, dense_torque_gap = 0, 0.0
    for k in contact:
        i, j = pr["p0"][k], pr["p1"][k]
        n = pr["nvec"][k]
        xju = x[i] + pr["r"][k] * n                    # particle j's image next to i
        crossing += abs(xju[2] - x[j][2]) > L[2] / 2
        R = pair_12(oracle, *oracle.dashpot_blocks(rad[i], rad[j], n, p))
        U = np.r_[u_inf(x[i], p.shear_rate), p.shear_rate * OMEGA1, u_inf(xju, p.shear_rate), p.shear_rate * OMEGA1]
        f = -R @ U
        F[i, :3] += f[:3]
        F[j, :3] += f[6:9]
        F[i, 3:] += rad[i] * np.cross(n, f[:3])
        F[j, 3:] += rad[j] * np.cross(n, f[:3])
        dense_torque_gap = max(dense_torque_gap, np.abs(f[3:6] - rad[i] * np.cross(n, f[:3])).max())
    assert crossing > 0                                # some contacts span the Lees-Edwards face
    assert np.abs(F[:, :3].sum(0)).max() <= 1e-12 * np.abs(F).max()       # pair forces cancel
    got = S.rhs_component("dashpot")
    assert np.abs(got - F).max() <= 1e-11 * np.abs(F).max()
    assert dense_torque_gap > 1e-3 * np.abs(F[:, 3:]).max()        # the two torque conventions differ
