"""Pins of the oracle (CPU restatement of LF_DEM's step path) against the
reference's own fixtures and against what the mathematics fixes.

Golden: LF_DEM/check/reference/int_*_repulsive_stress_stress5r.dat (t=0 pair
list, normal vectors and gaps of check/confs/D3N200VF0.45Bidi1.4_0.5Cubic_1.dat,
printed to 6 significant digits).  The rate-controlled golden of the same
suite is not used: its t=0 positions are not those of the .bin it names
(the suite is stale against the current source).
"""
import math

import numpy as np
import pytest

from tests.conftest import conf


def _pairs(d):
    return list(zip(d["p0"].tolist(), d["p1"].tolist()))


# ------------------------------------------------------------------ golden vectors
def test_golden_pair_set(golden, oracle):
    S = oracle.System(golden["dat_pos"], golden["dat_radius"], golden["dat_L"], golden["dat_disp"])
    pr = S.pairs()
    gi = golden["stress_int"]
    gp = set(zip(gi[:, 0].astype(int).tolist(), gi[:, 1].astype(int).tolist()))
    assert len(gp) == 745
    assert set(_pairs(pr)) == gp
    # pairs listed sorted by (p0, p1), as the reference prints them
    assert _pairs(pr) == sorted(gp)


def test_golden_geometry(golden, oracle):
    S = oracle.System(golden["dat_pos"], golden["dat_radius"], golden["dat_L"], golden["dat_disp"])
    pr = S.pairs()
    idx = {k: i for i, k in enumerate(_pairs(pr))}
    gi = golden["stress_int"]
    k = np.array([idx[(int(a), int(b))] for a, b in gi[:, :2]])
    # 6 significant digits in the file
    assert np.abs(pr["nvec"][k] - gi[:, 3:6]).max() <= 5.1e-7
    assert np.abs(pr["gap"][k] - gi[:, 6]).max() <= 5.1e-7 * np.maximum(1, np.abs(gi[:, 6])).max()
    # no contact at t=0 (column 3 = contact state 0)
    assert np.all(gi[:, 2] == 0) and np.all(pr["state"] == 0)


def test_golden_pair_count_in_data_file(golden, oracle):
    """data_* column 22 'number of interaction' is 745 in both goldens."""
    S = oracle.System(golden["dat_pos"], golden["dat_radius"], golden["dat_L"], golden["dat_disp"])
    assert S.npairs == 745


# ------------------------------------------------------------------ pair set
@pytest.mark.parametrize("case", ["dat", "bin"])
def test_cell_list_equals_bruteforce(golden, oracle, case):
    pos, rad, L, sd = (golden[case + k] for k in ("_pos", "_radius", "_L", "_disp"))
    S = oracle.System(pos, rad, L, sd)
    b0, b1 = oracle.bruteforce_pairs(pos, rad, L, sd)
    assert _pairs(S.pairs()) == list(zip(b0.tolist(), b1.tolist()))


def test_cell_list_boxed_c1(oracle):
    pos, rad, L = conf(500, 0.5, 1)
    for sd in ((0, 0, 0), (3.7, 0, 0), (L[0] - 1e-3, 0, 0)):
        S = oracle.System(pos, rad, L, sd)
        info = S.box_info()
        assert info["boxed"] and info["nx"] == 5
        b0, b1 = oracle.bruteforce_pairs(pos, rad, L, sd)
        assert _pairs(S.pairs()) == list(zip(b0.tolist(), b1.tolist()))


def test_pair_set_translation_invariant(oracle):
    """Shifting every particle by (dx, dy, dz) and re-wrapping through the
    Lees-Edwards boundary is the same suspension: same pairs, except for
    pairs whose distance is within rounding of the cutoff."""
    pos, rad, L = conf(500, 0.5, 2)
    sd = np.array([2.3, 0, 0])
    b0, b1 = oracle.bruteforce_pairs(pos, rad, L, sd)
    shifted = pos + np.array([1.7, -2.9, 4.1])
    for i in range(len(rad)):
        p = shifted[i].copy()
        oracle.lib().orc_periodize(oracle.P(L), oracle.P(sd), p.ctypes.data_as(oracle._dp))
        shifted[i] = p
    c0, c1 = oracle.bruteforce_pairs(shifted, rad, L, sd)
    A = set(zip(b0.tolist(), b1.tolist()))
    B = set(zip(c0.tolist(), c1.tolist()))
    for (i, j) in A ^ B:
        _, d = oracle.periodize_diff(L, sd, pos[j] - pos[i])
        rng = oracle.lib().orc_lub_range(rad[i], rad[j], 0.5)
        assert abs(np.dot(d, d) - rng * rng) < 1e-9
    assert len(A ^ B) <= 2


def test_periodize_diff_minimum_image(oracle):
    L = np.array([10.0, 12.0, 14.0])
    sd = np.array([3.0, 0.0, 0.0])
    z, d = oracle.periodize_diff(L, sd, [0.5, 0.5, 8.0])     # crosses the top: shift by -shear_disp
    assert z == -1 and np.allclose(d, [0.5 - 3.0, 0.5, 8.0 - 14.0])
    z, d = oracle.periodize_diff(L, sd, [-4.0, 0.5, 8.0])    # and the x wrap after the shift
    assert z == -1 and np.allclose(d, [-4.0 - 3.0 + 10.0, 0.5, -6.0])
    z, d = oracle.periodize_diff(L, sd, [0.5, -7.0, -9.0])
    assert z == 1 and np.allclose(d, [3.5, 5.0, 5.0])


def test_lub_range(oracle):
    r = oracle.lib().orc_lub_range
    assert r(1.0, 1.4, 0.5) == pytest.approx(2.5 * 1.2)
    assert r(1.0, 3.0, 0.5) == pytest.approx(4.0 + 0.5)   # ratio outside (0.5, 2)


# ------------------------------------------------------------------ resistance functions
def test_lub_constants_symmetry(oracle):
    """X_ab(l) = X_{3-a,3-b}(1/l) (XA, YA, YC, XM, YM: pair swap is the index
    reflection) and Y_ab(l) = -Y_{3-a,3-b}(1/l) (YB, XG, YG)
    (Jeffrey & Onishi 1984; Lubrication.cpp:190-293)."""
    for a0, a1 in ((1.0, 1.4), (1.4, 1.0), (1.0, 1.0), (0.7, 1.9)):
        c = oracle.lub_constants(a0, a1)
        s = oracle.lub_constants(a1, a0)
        for k in (0, 1, 3, 7, 8):          # XA YA YC XM YM
            assert c[k, 0] == pytest.approx(s[k, 3], rel=1e-14)
            assert c[k, 3] == pytest.approx(s[k, 0], rel=1e-14)
            assert c[k, 1] == pytest.approx(s[k, 2], rel=1e-14)
        for k in (2, 4, 5):                # YB XG YG
            assert c[k, 0] == pytest.approx(-s[k, 3], rel=1e-14)
            assert c[k, 1] == pytest.approx(-s[k, 2], rel=1e-14)


def test_lub_constants_equal_spheres(oracle):
    """Equal unit spheres: g1(XA)=1/4, g2(YA)=1/6, g2(YB)=-1/4, g2(YC)=1/5, g4(YC)=1/20."""
    c = oracle.lub_constants(1.0, 1.0)
    assert c[0, 0] == pytest.approx(0.25) and c[0, 1] == pytest.approx(-0.25)
    assert c[1, 0] == pytest.approx(1 / 6) and c[1, 1] == pytest.approx(-1 / 6)
    assert c[2, 0] == pytest.approx(2 / 3 * -0.25)
    assert c[3, 0] == pytest.approx(4 / 3 * 0.2) and c[3, 1] == pytest.approx(8 / 6 * 0.05)


def _pair_matrix(oracle, db0, db1, od):
    R = np.zeros((12, 12))
    R[:6, :6] = oracle.dblock_dense(db0)
    R[6:, 6:] = oracle.dblock_dense(db1)
    M = oracle.odblock_dense(od)
    R[6:, :6] = M
    R[:6, 6:] = M.T
    return R


NVECS = [np.array(v) / np.linalg.norm(v) for v in ([0.3, -0.5, 0.8], [1, 0, 0], [0.1, 0.9, -0.2])]


@pytest.mark.parametrize("kind", ["lub", "dashpot"])
@pytest.mark.parametrize("radii", [(1.0, 1.0), (1.0, 1.4), (1.4, 1.0)])
def test_pair_resistance_properties(oracle, kind, radii):
    p = oracle.default_params()
    a0, a1 = radii
    for n in NVECS:
        if kind == "lub":
            blocks = oracle.lub_blocks(a0, a1, 0.01, n, p)
            swapped = oracle.lub_blocks(a1, a0, 0.01, -n, p)
        else:
            blocks = oracle.dashpot_blocks(a0, a1, n, p)
            swapped = oracle.dashpot_blocks(a1, a0, -n, p)
        R = _pair_matrix(oracle, *blocks)
        # symmetric and positive semi-definite (dissipation >= 0)
        assert np.abs(R - R.T).max() == 0
        assert np.linalg.eigvalsh(R).min() >= -1e-12 * np.abs(R).max()
        # relabelling the pair (swap particles, flip n) is the same operator
        Rs = _pair_matrix(oracle, *swapped)
        P = np.zeros((12, 12))
        P[:6, 6:] = np.eye(6)
        P[6:, :6] = np.eye(6)
        assert np.abs(P @ Rs @ P - R).max() <= 1e-12 * np.abs(R).max()
        # rigid translation: no force from the singular squeeze/shear terms
        U = np.r_[1.0, -2.0, 0.5, 0, 0, 0, 1.0, -2.0, 0.5, 0, 0, 0]
        f = R @ U
        assert np.abs(f[[0, 1, 2, 6, 7, 8]]).max() <= 1e-12 * np.abs(R).max()
        # normal squeeze: force along n, magnitude XA11 - XA12 times relative speed
        Un = np.r_[n, 0, 0, 0, -n, 0, 0, 0]
        fn = R @ Un
        assert np.abs(np.cross(fn[:3], n)).max() <= 1e-12 * np.abs(R).max()


def test_lub_squeeze_matches_jeffrey_onishi(oracle):
    """Head-on approach of two unit spheres at gap h: force = 2 g1 / (h + delta) * U
    with g1 = 1/4 at lambda = 1 (the leading XA term), in LF_DEM's units."""
    p = oracle.default_params()
    n = np.array([0.0, 0.0, 1.0])
    for h in (1e-3, 1e-2, 0.1):
        db0, db1, od = oracle.lub_blocks(1.0, 1.0, h, n, p)
        R = _pair_matrix(oracle, db0, db1, od)
        U = np.r_[0, 0, 1.0, 0, 0, 0, 0, 0, -1.0, 0, 0, 0]
        F = R @ U
        assert F[2] == pytest.approx(2 * 0.25 / (h + p.lub_reduce), rel=1e-14)


def test_dashpot_coefficients(oracle):
    """Normal dashpot coefficient 4 kn tau_n, tangential log(1/delta)
    (ContactDashpot.cpp:65-105) on the XA11 / YA11 slots of equal spheres."""
    p = oracle.default_params()
    n = np.array([0.0, 0.0, 1.0])
    db0, _, _ = oracle.dashpot_blocks(1.0, 1.0, n, p)
    D = oracle.dblock_dense(db0)
    assert D[2, 2] == pytest.approx(4 * p.kn * p.tau_n * 0.25)
    assert D[0, 0] == pytest.approx(math.log(1 / p.lub_reduce) / 6)


# ------------------------------------------------------------------ matrix layout
def test_csc_export_matches_dense(oracle):
    """CSC lower triangle (StokesSolver.cpp:334-379) rebuilds the same
    symmetric matrix as the dense block expansion."""
    pos, rad, L = conf(500, 0.5, 1)
    S = oracle.System(pos, rad, L)
    S.first_update()
    db, cp, rows, od = S.blocks()
    n = len(rad)
    A = oracle.assemble_dense(n, db, cp, rows, od)
    nnz = oracle.lib().orc_csc_nnz(n, len(rows))
    colp = np.zeros(6 * n + 1, dtype=np.int32)
    rowi = np.zeros(nnz, dtype=np.int32)
    val = np.zeros(nnz)
    oracle.lib().orc_export_csc(n, oracle.P(db), len(rows), oracle.P(cp, np.int32), oracle.P(rows, np.int32),
                                oracle.P(od), colp.ctypes.data_as(oracle._ip), rowi.ctypes.data_as(oracle._ip),
                                val.ctypes.data_as(oracle._dp))
    assert colp[-1] == nnz
    B = np.zeros_like(A)
    for c in range(6 * n):
        for k in range(colp[c], colp[c + 1]):
            r = rowi[k]
            assert r >= c
            B[r, c] = val[k]
            B[c, r] = val[k]
        assert np.all(np.diff(rowi[colp[c]:colp[c + 1]]) > 0)    # sorted, as declared to CHOLMOD
    assert np.array_equal(A, B)


def test_cholesky_matches_dense_solve(oracle):
    pos, rad, L = conf(500, 0.5, 1)
    S = oracle.System(pos, rad, L)
    S.first_update()
    db, cp, rows, od = S.blocks()
    A = oracle.assemble_dense(len(rad), db, cp, rows, od)
    b = S.rhs()
    C = oracle.Chol(len(rad), db, cp, rows, od)
    assert C.status == 0
    x = C.solve(b)
    xd = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) <= 1e-12
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1e-13


def test_drag_only_limit(oracle):
    """Particles far apart: R is the drag diagonal, U = F / (sd a), W = T / (sd 4/3 a^3)."""
    pos = np.array([[1.0, 1.0, 1.0], [10.0, 10.0, 10.0], [1.0, 10.0, 5.0]])
    rad = np.array([1.0, 1.4, 1.0])
    S = oracle.System(pos, rad, [20.0, 20.0, 20.0])
    assert S.npairs == 0
    db, cp, rows, od = S.blocks()
    C = oracle.Chol(3, db, cp, rows, od)
    b = np.arange(1.0, 19.0)
    x = C.solve(b).reshape(3, 6)
    for i, a in enumerate(rad):
        assert np.allclose(x[i, :3], b.reshape(3, 6)[i, :3] / a, rtol=1e-15)
        assert np.allclose(x[i, 3:], b.reshape(3, 6)[i, 3:] / (4.0 / 3 * a ** 3), rtol=1e-15)


# ------------------------------------------------------------------ forces
def test_contact_spring_and_repulsion_closed_form(oracle):
    """Two overlapping spheres on the x axis: normal spring -kn (ro/2)^2 h
    (Contact.cpp:200-233) and repulsion F_R a0 a1/(a0+a1) at contact
    (RepulsiveForce.cpp:42-73) appear in the RHS with equal and opposite sign."""
    p = oracle.default_params(shear_rate=0.0)
    a0, a1 = 1.0, 1.4
    gap_abs = -0.01
    pos = np.array([[5.0, 5.0, 5.0], [5.0 + a0 + a1 + gap_abs, 5.0, 5.0]])
    S = oracle.System(pos, [a0, a1], [20.0, 20.0, 20.0], params=p)
    S.first_update()
    pr = S.pairs()
    assert pr["state"][0] == 2
    h = pr["gap"][0]
    rhs = S.rhs().reshape(2, 6)
    kn_s = ((a0 + a1) / 2) ** 2 * p.kn
    f_spring = kn_s * h                     # force on p0 along +x (n from 0 to 1)
    f_rep = -p.repulsion * a0 * a1 / (a0 + a1)
    assert rhs[0, 0] == pytest.approx(f_spring + f_rep, rel=1e-14)
    assert rhs[1, 0] == pytest.approx(-(f_spring + f_rep), rel=1e-14)
    assert np.abs(rhs[:, 1:3]).max() == 0 and np.abs(rhs[:, 3:]).max() == 0


def test_repulsion_decay(oracle):
    p = oracle.default_params(shear_rate=0.0)
    gap_abs = 0.05
    pos = np.array([[5.0, 5.0, 5.0], [5.0 + 2.0 + gap_abs, 5.0, 5.0]])
    S = oracle.System(pos, [1.0, 1.0], [20.0, 20.0, 20.0], params=p)
    S.first_update()
    rhs = S.rhs().reshape(2, 6)
    expect = -p.repulsion * 0.5 * math.exp(-gap_abs / p.repulsive_length)
    assert rhs[0, 0] == pytest.approx(expect, rel=1e-12)


def test_coulomb_slip_rescales_to_mu(oracle):
    """A contact sheared far past the static limit slides: |Ft| = mu |Fn|
    after the update (Contact.cpp:255-298)."""
    p = oracle.default_params(shear_rate=0.0, sd_coeff=1.0)
    pos = np.array([[5.0, 5.0, 5.0], [5.0 + 2.0 - 0.005, 5.0, 5.0]])
    S = oracle.System(pos, [1.0, 1.0], [20.0, 20.0, 20.0], params=p)
    S.first_update()
    S.set_pair_history([0], [1], disp_tan=[[0.0, 5.0, 0.0]])
    S.set_dt(0.0)
    S.set_in_predictor(True)     # the predictor keeps the stretch as prev_disp_tan
    S.update_interactions()
    pr = S.pairs()
    assert pr["state"][0] == 3
    kn_s, kt_s = p.kn, p.kt                  # ro/2 = 1
    fn = -kn_s * pr["gap"][0]
    ft = kt_s * np.linalg.norm(pr["disp_tan"][0])
    assert ft == pytest.approx(p.mu_dynamic * fn, rel=1e-12)


# ------------------------------------------------------------------ dynamics
def test_step_invariants(oracle):
    pos, rad, L = conf(500, 0.5, 3)
    S = oracle.System(pos, rad, L)
    S.first_update()
    strain0 = S.state()["strain"]
    for _ in range(3):
        S.step()
        st = S.state()
        assert 0 < st["dt"] <= S.params.disp_max / S.params.shear_rate + 1e-18
        assert np.all(st["pos"] >= 0) and np.all(st["pos"] < L)
        b0, b1 = oracle.bruteforce_pairs(st["pos"], rad, L, st["shear_disp"])
        assert set(zip(b0.tolist(), b1.tolist())) <= set(_pairs(S.pairs()))
    st = S.state()
    assert st["strain"] > strain0
    assert st["shear_disp"][0] == pytest.approx(st["strain"] * L[2] % L[0], rel=1e-9)


def test_strain_end_clamp(oracle):
    pos, rad, L = conf(500, 0.5, 3)
    S = oracle.System(pos, rad, L)
    S.first_update()
    end = 2.5e-4
    for _ in range(50):
        S.step(strain_end=end)
        s = S.state()["strain"]
        assert s <= end * (1 + 1e-12)
        if s > end - 1e-12:
            break
    assert S.state()["strain"] == pytest.approx(end, rel=1e-12)
