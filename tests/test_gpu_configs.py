"""GPU checks at the sizes of the benchmark workloads.

* N=4000 phi=0.56 (rate control, the headline workload): pair set, pair
  geometry and cell labels bit-exact against the oracle at t=0; velocities
  within 1e-8 of the oracle's FP64 sparse Cholesky solve at t=0, and again
  after 3 predictor-corrector steps (then on the device's own blocks and
  right-hand side, whose assembly is compared with the oracle's at t=0).
* N=32000 phi=0.58 (stress control, two right-hand sides per solve): t=0
  pair set and cells bit-exact against the oracle; blocks and right-hand
  side within FP64 rounding of the oracle's; FP64 residual of the device
  solution on the host below the solver target; after stress-controlled
  steps the shear stress equals the target and both right-hand sides per
  launch agree with one launch per right-hand side (the aggregate solver of
  large instances).
* 64-instance sweep (N in [250, 2000]): two steps of the whole batch, every
  instance bit-identical to its lone run; three instances against the
  oracle's pair sets and direct solve at t=0.
* served requests including one of N=8000: each request's final state
  equals its lone run.
"""
import numpy as np
import pytest

from tests.conftest import conf

pytestmark = pytest.mark.gpu

REL_VEL = 1e-8


def _pairset(d):
    return list(zip(d["p0"].tolist(), d["p1"].tolist()))


def _chol_of(oracle, n, db, od, p0, p1):
    cp = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(np.bincount(p0, minlength=n), out=cp[1:])
    return oracle.Chol(n, db, cp, p1.astype(np.int32), od)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _expansion(oracle):
    """Linear maps from LF_DEM's packed DBlock (15) / ODBlock (18) to the
    dense row-major 6x6 block (MatrixBlocks.h:5-84), from the oracle's
    expansion of unit blocks."""
    D = np.array([oracle.dblock_dense(np.eye(15)[k]).ravel() for k in range(15)])
    O = np.array([oracle.odblock_dense(np.eye(18)[k]).ravel() for k in range(18)])
    return D, O


def _host_residual(oracle, n, db, od, p0, p1, x, b):
    """||b - A x|| / ||b|| in FP64 on the host, A from the packed blocks."""
    D, O = _expansion(oracle)
    Dd = (db @ D).reshape(n, 6, 6)
    Od = (od @ O).reshape(-1, 6, 6)                 # block at (row p1, column p0)
    X = x.reshape(n, 6)
    Y = np.einsum("nij,nj->ni", Dd, X)
    np.add.at(Y, p1, np.einsum("kij,kj->ki", Od, X[p0]))
    np.add.at(Y, p0, np.einsum("kji,kj->ki", Od, X[p1]))
    return np.linalg.norm(b.reshape(n, 6) - Y) / np.linalg.norm(b)


# ---------------------------------------------------------------- N=4000, phi=0.56
def test_n4000_parity_at_t0_and_after_steps(capi, oracle):
    pos, rad, L = conf(4000, 0.56, 1)
    print("N=4000 phi=0.56: t=0 pairs and cells", flush=True)
    S = capi.System(pos, rad, L)
    R = oracle.System(pos, rad, L)
    g, o = S.pairs(), R.pairs()
    assert _pairset(g) == _pairset(o)
    for k in ("nvec", "r", "gap", "state"):
        assert np.array_equal(g[k], o[k]), k
    c, bi = S.cells(), R.box_info()
    assert (c["nx"], c["ny"], c["nz"]) == (bi["nx"], bi["ny"], bi["nz"])
    assert np.array_equal(c["labels"], bi["labels"])
    S.first_update()
    R.first_update()
    S.assemble()
    db, od, p0, p1 = S.blocks()
    rdb, rcp, rrows, rod = R.blocks()
    assert np.array_equal(np.repeat(np.arange(4000), np.diff(rcp)), p0) and np.array_equal(rrows, p1)
    assert np.abs(db - rdb).max() <= 1e-13 * np.abs(rdb).max()
    assert np.abs(od - rod).max() <= 1e-13 * np.abs(rod).max()
    print("N=4000: t=0 velocities against the oracle's direct solve", flush=True)
    S.compute_velocities(True)
    R.set_in_predictor(True)
    R.compute_velocities()
    gs, rs = S.state(), R.state()
    for k in ("na_vel", "na_ang_vel"):
        err = _rel(gs[k], rs[k])
        print(k, "rel err", err, flush=True)
        assert err <= REL_VEL
    print("N=4000: 3 steps, then the direct solve of the device's own system", flush=True)
    S.step(3)
    st = S.state()
    b0, b1 = oracle.bruteforce_pairs(st["pos"], rad, L, st["shear_disp"])
    assert set(zip(b0.tolist(), b1.tolist())) <= set(_pairset(S.pairs()))
    S.assemble()
    db, od, p0, p1 = S.blocks()
    b = S.rhs()
    S.compute_velocities(True)
    x = np.concatenate([S.state()["na_vel"], S.state()["na_ang_vel"]], axis=1).ravel()
    x_ref = _chol_of(oracle, 4000, db, od, p0, p1).solve(b)
    err = _rel(x, x_ref)
    print("after 3 steps: rel err", err, "contacts", int((S.pairs()["state"] >= 1).sum()), flush=True)
    assert err <= REL_VEL


# ---------------------------------------------------------------- N=32000, phi=0.58, stress control
@pytest.fixture(scope="module")
def c5():
    print("generating N=32000 phi=0.58", flush=True)
    return conf(32000, 0.58, 1)


def test_n32000_t0_pairs_blocks_and_host_residual(c5, capi, oracle):
    pos, rad, L = c5
    S = capi.System(pos, rad, L)
    R = oracle.System(pos, rad, L)
    print("N=32000: t=0 pairs", S.npairs, flush=True)
    g, o = S.pairs(), R.pairs()
    assert _pairset(g) == _pairset(o)
    assert np.array_equal(g["gap"], o["gap"]) and np.array_equal(g["nvec"], o["nvec"])
    assert np.array_equal(S.cells()["labels"], R.box_info()["labels"])
    S.first_update()
    R.first_update()
    S.assemble()
    db, od, p0, p1 = S.blocks()
    rdb, rcp, rrows, rod = R.blocks()
    assert np.array_equal(rrows, p1)
    assert np.abs(db - rdb).max() <= 1e-13 * np.abs(rdb).max()
    assert np.abs(od - rod).max() <= 1e-13 * np.abs(rod).max()
    b = S.rhs()
    assert np.abs(b - R.rhs()).max() <= 1e-13 * np.abs(b).max()
    print("N=32000: solve and FP64 residual on the host", flush=True)
    S.compute_velocities(True)
    st = S.solve_stats()
    x = np.concatenate([S.state()["na_vel"], S.state()["na_ang_vel"]], axis=1).ravel()
    res = _host_residual(oracle, 32000, db, od, p0, p1, x, b)
    print("host residual", res, "device", st, flush=True)
    assert res <= 2e-12


def test_n32000_stress_control(c5, capi):
    pos, rad, L = c5
    out = []
    for m in (0, 1):
        S = capi.System(pos, rad, L, opts=capi.default_opts(max_rhs=m))
        S.first_update()
        if m == 0:
            S.compute_stress(False)
            target = S.stress()["total"][2]
        S.set_control("stress", target)
        out.append(S)
    P, Q = out
    for k in range(2):
        print("N=32000 stress-controlled step", k, flush=True)
        for s in (P, Q):
            s.step(1)
            # the aggregate preconditioner of large instances (chunked grouping, 16-wave solver)
            assert s.precond_waves() > 0
            st = s.solve_stats()
            assert st["res_norm"] <= 1e-12 * st["rhs_norm"] * 1.0000001
        a, b = P.state(), Q.state()
        for key in ("na_vel", "na_ang_vel"):
            assert np.abs(a[key] - b[key]).max() <= 1e-9 * np.abs(b[key]).max(), key
        assert abs(a["strain"] - b["strain"]) <= 1e-9 * b["strain"]
    P.compute_stress(False)
    s = P.stress()
    print("shear rate", s["shear_rate"], "stress", s["total"][2], "target", target, flush=True)
    assert s["total"][2] == pytest.approx(target, rel=1e-8)


# ---------------------------------------------------------------- 64-instance sweep
def test_sweep_of_64_instances(capi, oracle):
    from lf_dem_amd import workloads
    specs = workloads.sweep_specs(64, 1)
    print("generating the 64-instance sweep, total N", sum(s[0] for s in specs), flush=True)
    confs = workloads.generate_many(specs, workers=8)
    params = [capi.default_params(repulsion=s[3]) for s in specs]
    opts = capi.default_opts(precond=0)
    slots = cells = 0
    for (pos, rad, L), p in zip(confs, params):
        sl, ce, _ = capi.footprint(len(rad), rad, L, p)
        slots += sl
        cells += ce
    B = capi.Batch(slots, cells, len(specs), opts=opts)
    ids = [B.add(pos, rad, L, params=p) for (pos, rad, L), p in zip(confs, params)]
    print("batch", B.info(), flush=True)
    B.step(2)
    B.sync()
    for k, ((pos, rad, L), p) in enumerate(zip(confs, params)):
        S = capi.System(pos, rad, L, params=p, opts=opts)
        S.first_update()
        S.step(2)
        a, b = B.state(ids[k]), S.state()
        for key in ("pos", "vel", "ang_vel", "na_vel", "na_ang_vel", "shear_disp"):
            assert np.array_equal(a[key], b[key]), (k, key)
        assert (a["dt"], a["strain"]) == (b["dt"], b["strain"]), k
        del S
    small = sorted(range(64), key=lambda k: specs[k][0])[:3]
    for k in small:
        pos, rad, L = confs[k]
        print("instance", k, "N", len(rad), "against the oracle at t=0", flush=True)
        S = capi.System(pos, rad, L, params=params[k], opts=opts)
        R = oracle.System(pos, rad, L, params=oracle.default_params(repulsion=specs[k][3]))
        assert _pairset(S.pairs()) == _pairset(R.pairs())
        S.first_update()
        R.first_update()
        S.compute_velocities(True)
        R.set_in_predictor(True)
        R.compute_velocities()
        for key in ("na_vel", "na_ang_vel"):
            assert _rel(S.state()[key], R.state()[key]) <= REL_VEL


# ---------------------------------------------------------------- serving with an N=8000 request
def test_served_mix_with_a_large_request(capi):
    from lf_dem_amd import serve
    specs = [(8000, 0.55, 31, 0.004), (600, 0.52, 32, 0.001), (1500, 0.57, 33, 0.002)]
    reqs = []
    for rid, (n, phi, seed, strain) in enumerate(specs):
        pos, rad, L = conf(n, phi, seed)
        reqs.append(serve.Request(rid, pos, rad, L, strain=strain))
    # room for the first two only: the third waits for the second to retire
    srv = serve.Server(reqs[0].slots + reqs[1].slots, reqs[0].cells + reqs[1].cells, 4)
    for r in reqs:
        srv.submit(r)
    k = 0
    while srv.live or srv.waiting:
        srv.admit(k)
        srv.step()
        k += 1
        assert k < 2000
    assert reqs[2].admitted > 0 and all(r.status == 0 for r in reqs)
    for r in reqs:
        print("request", r.rid, "N", len(r.radius), "steps", r.steps, flush=True)
        S = capi.System(r.pos, r.radius, r.L, params=r.params)
        S.first_update()
        steps = 0
        while True:
            S.step(1, r.strain)
            steps += 1
            if S.state()["strain"] > r.strain - 1e-8:
                break
        assert steps == r.steps
        a, b = r.state, S.state()
        for key in ("pos", "vel", "ang_vel", "na_vel", "na_ang_vel", "shear_disp"):
            assert np.array_equal(a[key], b[key]), (r.rid, key)
