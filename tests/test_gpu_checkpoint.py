"""Checkpoint and restart of a device system (Simulation::checkpoint,
System::setContacts / getContacts) through the C ABI.

The restart restores what LF_DEM's checkpoint holds: positions, Lees-Edwards
displacement and the clock bit for bit, and the active contacts' tangential
displacements, which then pass one pair update with dt = 0 (LF_DEM's first
update after setContacts) and so agree to rounding.  Velocities and the time step are recomputed from them, as
in LF_DEM, so the continued trajectory stays close to, not equal to, the
uninterrupted one.
"""
import numpy as np
import pytest

from tests.conftest import conf

pytestmark = pytest.mark.gpu


def test_checkpoint_restart(capi, tmp_path):
    from lf_dem_amd import checkpoint
    pos, rad, L = conf(500, 0.55, 3)
    params = capi.default_params(repulsion=0.05)
    S = capi.System(pos, rad, L, params=params)
    S.first_update()
    for k in range(60):
        S.step(1)
        c = S.contacts()
        if k >= 9 and len(c["p0"]) >= 5 and np.abs(c["disp_tan"]).max() > 0:
            break
    print("steps", k + 1, "contacts", len(c["p0"]), flush=True)
    assert len(c["p0"]) >= 5 and np.abs(c["disp_tan"]).max() > 0
    conf_path, chk_path = checkpoint.checkpoint(S, "t", str(tmp_path))
    R, clock = checkpoint.resume(conf_path, chk_path, params=params)
    a, b = S.state(), R.state()
    assert np.array_equal(a["pos"], b["pos"]) and np.array_equal(a["shear_disp"], b["shear_disp"])
    assert (b["time"], b["strain"]) == (a["time"], a["strain"]) == (clock.time, clock.cumulated_strain)
    ca, cb = S.contacts(), R.contacts()
    assert np.array_equal(ca["p0"], cb["p0"]) and np.array_equal(ca["p1"], cb["p1"])
    # restored exactly, then one dt = 0 pair update (tangent-plane projection,
    # Coulomb check), as LF_DEM's first update after setContacts: rounding level
    dev = np.abs(ca["disp_tan"] - cb["disp_tan"]).max()
    print("disp_tan deviation after the restart update", dev, "scale", np.abs(ca["disp_tan"]).max(), flush=True)
    assert dev <= 1e-12 * np.abs(ca["disp_tan"]).max()
    pa, pb = S.pairs(), R.pairs()
    assert np.array_equal(pa["p0"], pb["p0"]) and np.array_equal(pa["p1"], pb["p1"])
    # the continued runs: the restart recomputes velocities and the time step,
    # so it deviates from the uninterrupted run well within that run's own
    # motion over the same steps
    p0 = S.state()["pos"]
    S.step(3)
    R.step(3)
    p3 = S.state()["pos"]
    d = p3 - R.state()["pos"]
    d -= np.round(d / L) * L
    m = p3 - p0
    m -= np.round(m / L) * L
    print("position deviation after 3 steps", np.abs(d).max(), "motion", np.abs(m).max(), flush=True)
    assert np.abs(d).max() <= 0.5 * np.abs(m).max()


def test_set_contacts_rejects_unknown_pairs(capi):
    pos, rad, L = conf(300, 0.5, 2)
    S = capi.System(pos, rad, L)
    S.first_update()
    p = S.pairs()
    with pytest.raises(capi.LfdError):
        S.set_contacts([0], [0], np.zeros((1, 3)))            # no interaction (0, 0)
    with pytest.raises(capi.LfdError):
        S.set_contacts([p["p0"][0]], [p["p1"][0]], np.zeros((1, 3)), np.ones((1, 3)))   # rolling friction
