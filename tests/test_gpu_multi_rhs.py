"""Paired right-hand sides in one launch of the LDS-staged block-Jacobi
solver (lfd_solve_opts.max_rhs = 0) against one launch per right-hand side
(max_rhs = 1), through the C ABI.

Both forms apply the same operations in the same order to each right-hand
side (one that stops early is frozen while the other iterates), but the grid
sums carry each round's tag in the low 8 mantissa bits of their partial
records and the two forms publish different round sequences, so the sums
differ at the 2^-44 level.  The bar is agreement to the solve tolerance:
velocities within 1e-9 of their scale, positions within 1e-9, over
stress-controlled steps (two right-hand sides per velocity computation), the
stress evaluation (four), a zero right-hand side beside a nonzero one,
captured steps and a batch of two instances.
"""
import numpy as np
import pytest

from tests.conftest import conf

pytestmark = pytest.mark.gpu


def _same(a, b):
    for k in ("vel", "ang_vel", "na_vel", "na_ang_vel"):
        scale = max(np.abs(b[k]).max(), 1e-300)
        assert np.abs(a[k]-b[k]).max() <= 1e-9*scale, (k, np.abs(a[k]-b[k]).max()/scale)
    for k in ("pos", "shear_disp"):
        assert np.abs(np.asarray(a[k])-np.asarray(b[k])).max() <= 1e-9, (k, np.abs(np.asarray(a[k])-np.asarray(b[k])).max())
    for k in ("dt", "time", "strain"):
        assert abs(a[k]-b[k]) <= 1e-9*abs(b[k])+1e-300, k


def _pair(capi, pos, rad, L, params=None, launch=0):
    out = []
    for m in (0, 1):
        s = capi.System(pos, rad, L, params=params, opts=capi.default_opts(precond=0, max_rhs=m, launch=launch))
        s.first_update()
        out.append(s)
    out[0].compute_stress(False)
    target = out[0].stress()["total"][2]
    for s in out:
        s.set_control("stress", target)
    return out


@pytest.mark.parametrize("n,phi,seed", [(600, 0.52, 3), (1500, 0.56, 4)])
def test_stress_control_paired_matches_single(n, phi, seed, capi):
    pos, rad, L = conf(n, phi, seed)
    P, S = _pair(capi, pos, rad, L)
    launches = []
    for s in (P, S):
        s.profile(True)
        s.step(1)
        launches.append(s.profile(False)[1])
    # two velocity computations per step, two right-hand sides each
    assert launches == [2, 4]
    _same(P.state(), S.state())
    for k in range(5):
        print("n", n, "stress-controlled step", k, flush=True)
        P.step(1)
        S.step(1)
        _same(P.state(), S.state())
    print("n", n, "stress evaluation", flush=True)
    P.step_stress()
    S.step_stress()
    _same(P.state(), S.state())
    a, b = P.stress(), S.stress()
    assert np.allclose(a["total"], b["total"], rtol=1e-8, atol=1e-12*np.abs(b["total"]).max())
    assert abs(a["shear_rate"]-b["shear_rate"]) <= 1e-9*abs(b["shear_rate"])
    print("n", n, "captured steps", flush=True)
    P.step(3, graph=True)
    S.step(3, graph=True)
    _same(P.state(), S.state())


def test_zero_right_hand_side_beside_a_nonzero_one(capi):
    """Dilute and without repulsion: the rate independent right-hand side is
    zero (solution zero at once) while the rate proportional one iterates."""
    pos, rad, L = conf(500, 0.30, 5)
    P, S = _pair(capi, pos, rad, L, params=capi.default_params(repulsion=0.0))
    for k in range(4):
        print("dilute stress-controlled step", k, flush=True)
        P.step(1)
        S.step(1)
        _same(P.state(), S.state())
        # the paired launch reports the nonzero side's residual, not 0 / 0
        st = P.solve_stats()
        assert st["rhs_norm"] > 0 and st["res_norm"] <= 1e-12 * st["rhs_norm"] * 1.0000001
    assert np.isfinite(P.state()["vel"]).all()


def test_batch_paired_matches_single(capi):
    specs = [(400, 0.50, 21, 2.0), (900, 0.55, 22, 0.5)]
    batches = []
    for m in (0, 1):
        slots = cells = 0
        confs = []
        for n, phi, seed, rep in specs:
            pos, rad, L = conf(n, phi, seed)
            p = capi.default_params(repulsion=rep)
            sl, ce, _ = capi.footprint(len(rad), rad, L, p)
            slots += sl
            cells += ce
            confs.append((pos, rad, L, p))
        B = capi.Batch(slots, cells, len(specs), control="stress", opts=capi.default_opts(precond=0, max_rhs=m))
        ids = [B.add(pos, rad, L, params=p, target_stress=0.3) for pos, rad, L, p in confs]
        batches.append((B, ids))
    (P, ip), (S, is_) = batches
    for k in range(4):
        print("batch stress-controlled step", k, flush=True)
        P.step(1)
        S.step(1)
        P.sync()
        S.sync()
        for a, b in zip(ip, is_):
            _same(P.state(a), S.state(b))


def test_large_instance_paired_matches_single(capi):
    """An instance of 71 workgroups (N=9000): two stress-controlled steps,
    paired against one launch per right-hand side; both reach the FP64
    residual target."""
    pos, rad, L = conf(9000, 0.5, 5)
    P, S = _pair(capi, pos, rad, L)
    for k in range(2):
        print("N=9000 stress-controlled step", k, flush=True)
        P.step(1)
        S.step(1)
        for s in (P, S):
            st = s.solve_stats()
            assert st["res_norm"] <= 1e-12 * st["rhs_norm"] * 1.0000001
        _same(P.state(), S.state())


def test_paired_cooperative_launch(capi):
    """Both forms launched cooperatively (as at N=32000, where the grid takes
    nearly every CU), against each other and against the plain launch."""
    pos, rad, L = conf(1500, 0.56, 4)
    P, S = _pair(capi, pos, rad, L, launch=1)
    Q, _ = _pair(capi, pos, rad, L, launch=2)
    for k in range(3):
        print("cooperative stress-controlled step", k, flush=True)
        for s in (P, S, Q):
            s.step(1)
        _same(P.state(), S.state())
        _same(P.state(), Q.state())
