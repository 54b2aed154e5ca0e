"""Interaction snapshot forces and non-affine displacement on the GPU
(lfd_system_get_pair_output / lfd_system_get_na_disp through the C ABI)
against the oracle, whose columns are pinned by force balance in
test_oracle_pair_output.py.  Bars: every column within 1e-8 of its largest
magnitude over the pairs; na_disp within 1e-8 of its largest magnitude
(both follow the velocities, which agree to the solve tolerance)."""
import numpy as np
import pytest

from tests.conftest import conf

pytestmark = pytest.mark.gpu

REL = 1e-8


@pytest.fixture(scope="module")
def stepped(capi, oracle):
    pos, rad, L = conf(500, 0.5, 1)
    S = capi.System(pos, rad, L)
    R = oracle.System(pos, rad, L)
    S.first_update()
    R.first_update()
    for _ in range(3):
        S.step()
        R.step()
    return S, R, (pos, rad, L)


def _cmp_cols(g, o):
    assert g.shape == o.shape
    for c in range(o.shape[1]):
        scale = max(np.abs(o[:, c]).max(), 1e-300)
        assert np.abs(g[:, c]-o[:, c]).max() <= REL*scale, c


def test_pair_output_after_steps(stepped):
    S, R, _ = stepped
    gp, op = S.pairs(), R.pairs()
    assert np.array_equal(gp["p0"], op["p0"]) and np.array_equal(gp["p1"], op["p1"])
    assert (gp["state"] > 0).sum() > 100
    print("pairs", len(gp["p0"]), "contacts", (gp["state"] > 0).sum(), flush=True)
    _cmp_cols(S.pair_output(), R.pair_output())


def test_na_disp_after_steps(stepped):
    S, R, _ = stepped
    g, o = S.na_disp(), R.na_disp()
    assert np.abs(o).max() > 0
    assert np.abs(g-o).max() <= REL*np.abs(o).max()


def test_pair_forces_balance_on_gpu(stepped):
    """The device's columns balance the Stokes drag at the velocities of the
    solve that produced them, to the solve tolerance."""
    S, R, (pos, rad, L) = stepped
    S.compute_velocities(False)
    po, pr, st = S.pair_output(), S.pairs(), S.state()
    n = pr["nvec"]
    tot = -po[:, [0]]*n + po[:, 1:4] + po[:, [4]]*n + po[:, 5:8] - po[:, [8]]*n
    F = np.zeros((len(rad), 3))
    np.add.at(F, pr["p0"], tot)
    np.add.at(F, pr["p1"], -tot)
    drag = -rad[:, None]*st["na_vel"]
    assert np.abs(F+drag).max() <= 1e-8*np.abs(drag).max()


def test_batch_instance_pair_output(capi, stepped):
    """The batch form on an instance that is not the first of its batch."""
    _, _, (pos, rad, L) = stepped
    p2, r2, L2 = conf(300, 0.5, 2)
    slots = sum(capi.footprint(len(r), r, l)[0] for r, l in ((r2, L2), (rad, L)))
    cells = sum(capi.footprint(len(r), r, l)[1] for r, l in ((r2, L2), (rad, L)))
    B = capi.Batch(slots, cells, 2)
    B.add(p2, r2, L2)
    iid = B.add(pos, rad, L)
    S = capi.System(pos, rad, L)
    S.first_update()
    B.step(2)
    S.step(2)
    np.testing.assert_array_equal(B.pair_output(iid), S.pair_output())
    np.testing.assert_array_equal(B.na_disp(iid), S.na_disp())
