"""Parity of the HIP path (through the C ABI) against the oracle.

Bars: pair set and cell labels bit-exact; pair geometry and assembled 6x6
blocks / RHS within a few FP64 ulps (the only differences allowed come from
the device log/exp, which need not round like glibc's); velocities within
1e-8 relative of the oracle's FP64 direct (Cholesky) solve.
"""
import numpy as np
import pytest

from tests.conftest import conf

pytestmark = pytest.mark.gpu

REL_VEL = 1e-8        # velocities vs FP64 direct solve
REL_BLOCK = 1e-13     # FP64 assembly: ulp-level (device log/exp vs glibc)


def _fixture_sys(golden):
    return golden["dat_pos"], golden["dat_radius"], golden["dat_L"], golden["dat_disp"]


def _bin_sys(golden):
    return golden["bin_pos"], golden["bin_radius"], golden["bin_L"], golden["bin_disp"]


def _c1():
    pos, rad, L = conf(500, 0.5, 1)
    return pos, rad, L, np.zeros(3)


CASES = {"n200_fixture": _fixture_sys, "n200_le_shift": _bin_sys}


def _make(case, golden):
    if case == "c1_n500":
        return _c1()
    return CASES[case](golden)


ALL = ["n200_fixture", "n200_le_shift", "c1_n500"]


def _pairset(d):
    return list(zip(d["p0"].tolist(), d["p1"].tolist()))


@pytest.mark.parametrize("case", ALL)
def test_pairs_and_cells_bitexact(case, golden, oracle, capi):
    pos, rad, L, sd = _make(case, golden)
    print("case", case, "np", len(rad), flush=True)
    S = capi.System(pos, rad, L, sd)
    R = oracle.System(pos, rad, L, sd)
    g, o = S.pairs(), R.pairs()
    assert _pairset(g) == _pairset(o)
    assert np.array_equal(g["nvec"], o["nvec"])
    assert np.array_equal(g["r"], o["r"])
    assert np.array_equal(g["gap"], o["gap"])
    assert np.array_equal(g["state"], o["state"])
    c, bi = S.cells(), R.box_info()
    assert (c["nx"], c["ny"], c["nz"], c["boxed"]) == (bi["nx"], bi["ny"], bi["nz"], bi["boxed"])
    assert np.array_equal(c["labels"], bi["labels"])


def test_golden_pair_set_on_gpu(golden, capi):
    """The reference's own t=0 pair list (check/reference/int_*stress5r.dat)."""
    pos, rad, L, sd = _fixture_sys(golden)
    S = capi.System(pos, rad, L, sd)
    g = S.pairs()
    gi = golden["stress_int"]
    assert set(_pairset(g)) == set(zip(gi[:, 0].astype(int).tolist(), gi[:, 1].astype(int).tolist()))
    assert len(g["p0"]) == 745


@pytest.mark.parametrize("case", ALL)
def test_blocks_and_rhs(case, golden, oracle, capi):
    pos, rad, L, sd = _make(case, golden)
    print("case", case, flush=True)
    S = capi.System(pos, rad, L, sd)
    R = oracle.System(pos, rad, L, sd)
    S.first_update()
    R.first_update()
    S.assemble()
    db, od, p0, p1 = S.blocks()
    rdb, rcp, rrows, rod = R.blocks()
    # every pair of these configurations carries a resistance
    assert len(rrows) == len(p1)
    rp0 = np.repeat(np.arange(len(rad)), np.diff(rcp))
    assert np.array_equal(rp0, p0) and np.array_equal(rrows, p1)
    scale = np.abs(rdb).max()
    assert np.abs(db - rdb).max() <= REL_BLOCK * scale
    assert np.abs(od - rod).max() <= REL_BLOCK * np.abs(rod).max()
    r_gpu, r_orc = S.rhs(), R.rhs()
    assert np.abs(r_gpu - r_orc).max() <= REL_BLOCK * np.abs(r_orc).max()


# (solver kind, preconditioner, aggregate waves per workgroup): mixed with
# the aggregate preconditioner (the default 8 waves per workgroup with the
# inverses in registers, and 12 with the inverses in LDS), mixed with 6x6
# block-Jacobi, plain FP64, mixed from HBM
SOLVERS = [(0, 1, 0), (0, 1, 12), (0, 0, 0), (1, 0, 0), (2, 0, 0)]


@pytest.mark.parametrize("kind,precond,agg_waves", SOLVERS)
@pytest.mark.parametrize("case", ALL)
def test_velocities_vs_direct(case, kind, precond, agg_waves, golden, oracle, capi):
    pos, rad, L, sd = _make(case, golden)
    print("case", case, "solver kind", kind, "preconditioner", precond, "waves", agg_waves, flush=True)
    opts = capi.default_opts(kind=kind, precond=precond, agg_waves=agg_waves)
    S = capi.System(pos, rad, L, sd, opts=opts)
    R = oracle.System(pos, rad, L, sd)
    S.first_update()
    R.first_update()
    S.compute_velocities(True)
    R.set_in_predictor(True)
    R.compute_velocities()
    gs, rs = S.state(), R.state()
    for key in ("na_vel", "na_ang_vel"):
        err = np.linalg.norm(gs[key] - rs[key]) / np.linalg.norm(rs[key])
        print(key, "rel err", err)
        assert err <= REL_VEL
    assert gs["max_velocity"] == pytest.approx(rs["max_velocity"], rel=1e-8)


@pytest.mark.parametrize("kind,precond", [(0, 1), (0, 0)])
def test_plain_and_cooperative_launch_agree(kind, precond, capi):
    """The persistent solvers give bit-identical steps whether launched
    cooperatively or plainly (the launch only decides residency, not the
    arithmetic)."""
    pos, rad, L, sd = _c1()
    out = []
    for launch in (1, 2):
        print("solver kind", kind, "preconditioner", precond, "launch", launch, flush=True)
        S = capi.System(pos, rad, L, sd, opts=capi.default_opts(kind=kind, precond=precond, launch=launch))
        S.first_update()
        S.step(3)
        out.append(S.state())
    for key in ("pos", "vel", "ang_vel", "na_vel"):
        assert np.array_equal(out[0][key], out[1][key]), key



def test_plain_form_with_device_scalars_equals_plain(capi):
    """LFD_SOLVER_PLAIN_DEVICE runs the FP64 PCG of LFD_SOLVER_PLAIN with its
    scalars on the device: the same iterates, so bit-identical steps and the
    same iteration counts, and the velocities within 1e-8 of the mixed form."""
    pos, rad, L, sd = _c1()
    out, its = [], []
    for kind in (1, 3, 0):
        S = capi.System(pos, rad, L, sd, opts=capi.default_opts(kind=kind, precond=0, tol=1e-10, max_inner=20000))
        S.first_update()
        S.step(3)
        out.append(S.state())
        its.append(S.solve_stats())
        print("kind", kind, its[-1], flush=True)
    for key in ("pos", "vel", "ang_vel", "na_vel", "na_ang_vel"):
        assert np.array_equal(out[0][key], out[1][key]), key
    assert its[0]["iters_inner"] == its[1]["iters_inner"] and its[0]["total_inner"] == its[1]["total_inner"]
    for key in ("na_vel", "na_ang_vel"):
        assert np.abs(out[1][key] - out[2][key]).max() <= 1e-8 * np.abs(out[2][key]).max(), key

def _pair_blocks(oracle, R, rad):
    """Per pair (p0, p1, DBlock of p0, DBlock of p1, ODBlock) as
    Interaction::RFU_DBlocks / RFU_ODBlock give them (Interaction.cpp:238-311):
    dashpot blocks in contact, lubrication blocks otherwise."""
    pr = R.pairs()
    out = []
    for k in range(len(pr["p0"])):
        i, j = int(pr["p0"][k]), int(pr["p1"][k])
        if pr["state"][k] >= 1:
            b = oracle.dashpot_blocks(rad[i], rad[j], pr["nvec"][k], R.params)
        else:
            b = oracle.lub_blocks(rad[i], rad[j], pr["gap"][k], pr["nvec"][k], R.params)
        out.append((i, j) + tuple(b))
    return out


def _drag(rad, sd_coeff):
    """The drag seed of System.cpp:553-575 as packed DBlocks."""
    d = np.zeros((len(rad), 15))
    d[:, [0, 5, 8]] = (sd_coeff * rad)[:, None]
    d[:, [9, 12, 14]] = (sd_coeff * 4.0 / 3 * rad ** 3)[:, None]
    return d


def test_stokes_dropin_matches_direct(golden, oracle, capi):
    """The system handed over through the StokesSolver column protocol as
    System::buildResistanceMatrix does (System.cpp:1429-1473): drag blocks
    at the reset, then per column every pair's two DBlocks and its ODBlock,
    the rows of a column in a shuffled order (StokesSolver.h:104-105 allows
    any order).  The diagonal accumulation must rebuild the oracle's
    assembled matrix; the solve must match its direct solve."""
    pos, rad, L, sd = _c1()
    R = oracle.System(pos, rad, L, sd)
    R.first_update()
    db, cp, rows, od = R.blocks()
    b = R.rhs()
    n = len(rad)
    blocks = _pair_blocks(oracle, R, rad)
    drag = _drag(rad, R.params.sd_coeff)
    acc = drag.copy()
    for i, j, di, dj, _ in blocks:
        acc[i] += di
        acc[j] += dj
    assert np.abs(acc - db).max() <= 1e-12 * np.abs(db).max()
    assert any(1 <= s for s in R.pairs()["state"])           # contacts: dashpot blocks take part
    ss = capi.StokesSolver(n)
    ss.resetResistanceMatrix(len(blocks), drag, True)
    rng = np.random.default_rng(5)
    by_col = {}
    for blk in blocks:
        by_col.setdefault(blk[0], []).append(blk)
    for i in range(n - 1):
        ss.startNewColumn()
        col = by_col.get(i, [])
        for t in rng.permutation(len(col)):
            _, j, di, dj, o = col[t]
            ss.addResistanceBlocks(i, j, di, dj, o)
    ss.matrixFillingDone()
    x_ref = oracle.Chol(n, db, cp, rows, od).solve(b)
    A = oracle.assemble_dense(n, db, cp, rows, od)
    y = ss.multiply(x_ref)
    assert np.abs(y - A @ x_ref).max() <= 1e-12 * np.abs(A @ x_ref).max()
    ss.resetRHS()
    ss.setRHS(b)
    u, w = ss.solve()
    x = np.concatenate([u, w], axis=1).ravel()
    err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    print("drop-in rel err", err, ss.last_stats.iters_inner, ss.last_stats.iters_outer)
    assert err <= REL_VEL
    with pytest.raises(capi.LfdError):
        ss.solve(out_len=6 * n - 1)   # StokesSolver::solve throws when the output is too small


def test_multi_step_trajectory(golden, oracle, capi):
    """A few predictor-corrector steps: same pair sets, positions within 1e-8."""
    pos, rad, L, sd = _c1()
    S = capi.System(pos, rad, L, sd)
    R = oracle.System(pos, rad, L, sd)
    S.first_update()
    R.first_update()
    for k in range(5):
        S.step(1)
        R.step()
        gs, rs = S.state(), R.state()
        print("step", k, "dt", gs["dt"], rs["dt"], flush=True)
        assert gs["dt"] == pytest.approx(rs["dt"], rel=1e-7)
        d = gs["pos"] - rs["pos"]
        d -= np.round(d / L) * L
        assert np.abs(d).max() <= 1e-8
        assert _pairset(S.pairs()) == _pairset(R.pairs())


def test_reference_rate_check_inputs_on_gpu(oracle, capi):
    """LF_DEM's rate-controlled check (`-r 0.2r confs/D3N200...bin
    repulsive.txt`, check/Makefile) read through lf_dem_amd.formats and
    lf_dem_amd.params: the GPU path follows the oracle over 5 steps (pair set
    bit-exact, positions within 1e-8)."""
    import os
    from lf_dem_amd import formats, params
    check = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lfdem_check")
    c = formats.read(os.path.join(check, "D3N200VF0.45Bidi1.4_0.5Cubic_1.bin"))
    P = params.load(os.path.join(check, "repulsive.txt"), "rate", "0.2r", legacy_keys=True)
    S = capi.System(c.pos, c.radius, c.L, c.shear_disp, params=capi.default_params(**P.params))
    R = oracle.System(c.pos, c.radius, c.L, c.shear_disp, params=oracle.default_params(**P.params))
    S.first_update()
    R.first_update()
    for _ in range(5):
        S.step(1)
        R.step()
    gp, rp = S.pairs(), R.pairs()
    assert np.array_equal(gp["p0"], rp["p0"]) and np.array_equal(gp["p1"], rp["p1"])
    g, o = S.state(), R.state()
    assert g["dt"] == pytest.approx(5e-5, rel=1e-12) and o["dt"] == pytest.approx(5e-5, rel=1e-12)
    d = g["pos"] - o["pos"]
    d -= np.round(d / c.L) * c.L
    assert np.abs(d).max() <= 1e-8


def test_large_instance_lds_and_hbm_solvers_agree(capi):
    """An instance of 71 workgroups (N=9000) in the LDS block-Jacobi solver
    (tagged partial records polled by every workgroup) and in the
    HBM-streamed solver (grid barrier): velocities agree to the solve
    tolerance and both reach the FP64 residual target."""
    pos, rad, L = conf(9000, 0.5, 5)
    out = []
    for kind in (0, 2):
        print("N=9000 solver kind", kind, flush=True)
        S = capi.System(pos, rad, L, opts=capi.default_opts(kind=kind, precond=0))
        S.first_update()
        S.compute_velocities(True)
        st = S.solve_stats()
        assert st["res_norm"] <= 1e-12 * st["rhs_norm"] * 1.0000001
        out.append(S.state())
    for key in ("na_vel", "na_ang_vel"):
        err = np.linalg.norm(out[0][key] - out[1][key]) / np.linalg.norm(out[1][key])
        print(key, "rel diff", err)
        assert err <= 1e-9
