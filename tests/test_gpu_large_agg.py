"""GPU checks of the aggregate preconditioner on instances above 8192
particles (chunked grouping, 16-wave solver with Eisenstat's product).

* the grouping is a valid packing: every particle one solver slot, at most 8
  particles per wave, the waves within the instance's budget, and no wave
  mixing particles of two chunks (cell-ordered positions, chunks of at most
  8192);
* the velocities reach the FP64 residual target (checked on the host from the
  device's own blocks) and equal the 6x6 block-Jacobi solver's to 1e-9; a few
  steps later the trajectories still agree;
* steps replayed from captured step graphs equal eager steps bit for bit;
* a large instance in a batch with a small one: each equal to its lone run;
* the same form asked for on small instances (agg_waves = 16): batch equal
  to lone runs, t=0 velocities within 1e-8 of the oracle's direct solve, two
  right-hand sides per launch equal to one per launch within 1e-9.
"""
import numpy as np
import pytest

from tests.conftest import conf
from tests.test_gpu_configs import _host_residual, _rel

pytestmark = pytest.mark.gpu

N, PHI = 9000, 0.50
CHUNK_MAX = 8192


@pytest.fixture(scope="module")
def c9k():
    print("generating N=%d phi=%.2f" % (N, PHI), flush=True)
    return conf(N, PHI, 1)


def _cell_order_chunks(S, n):
    """Chunk of each particle: positions in cell order (cells by label,
    members by index), split into ceil(n / 8192) chunks of equal size."""
    labels = S.cells()["labels"]
    order = np.lexsort((np.arange(n), labels))
    rank = np.empty(n, dtype=np.int64)
    rank[order] = np.arange(n)
    nch = -(-n // CHUNK_MAX)
    cs = -(-n // nch)
    return rank // cs


def test_grouping_is_a_valid_chunked_packing(c9k, capi):
    pos, rad, L = c9k
    S = capi.System(pos, rad, L, opts=capi.default_opts(precond=1))
    chunk = _cell_order_chunks(S, N)
    S.first_update()
    S.compute_velocities(True)
    waves = S.precond_waves()
    slots = S.solver_slots()
    print("waves", waves, "particles per wave %.3f" % (N / waves), flush=True)
    assert -(-N // 8) <= waves <= 16 * 8 * 256 // 8
    assert slots.min() >= 0 and slots.max() < 8 * waves
    assert len(np.unique(slots)) == N
    wave_of = slots // 8
    for w in np.unique(wave_of):
        assert len(np.unique(chunk[wave_of == w])) == 1, w
    # the strongest couplings are mostly inside waves: a third or more of the pairs
    pr = S.pairs()
    inside = np.count_nonzero(wave_of[pr["p0"]] == wave_of[pr["p1"]])
    print("pairs inside waves %d of %d" % (inside, len(pr["p0"])), flush=True)
    assert inside >= len(pr["p0"]) // 4


def test_velocities_against_block_jacobi_and_host_residual(c9k, capi, oracle):
    pos, rad, L = c9k
    out = {}
    for pc in (1, 0):
        S = capi.System(pos, rad, L, opts=capi.default_opts(precond=pc))
        S.first_update()
        S.assemble()
        db, od, p0, p1 = S.blocks()
        b = S.rhs()
        S.compute_velocities(True)
        st = S.solve_stats()
        x = np.concatenate([S.state()["na_vel"], S.state()["na_ang_vel"]], axis=1).ravel()
        res = _host_residual(oracle, N, db, od, p0, p1, x, b)
        print("precond", pc, "waves", S.precond_waves(), "host residual", res, st, flush=True)
        assert res <= 2e-12
        out[pc] = S
    for key in ("na_vel", "na_ang_vel"):
        assert _rel(out[1].state()[key], out[0].state()[key]) <= 1e-9, key
    for s in out.values():
        s.step(5)                      # through a regrouping (every 4 steps)
    a, b = out[1].state(), out[0].state()
    assert np.abs(a["pos"] - b["pos"]).max() <= 1e-8 * L.max()
    for key in ("na_vel", "na_ang_vel"):
        assert _rel(a[key], b[key]) <= 1e-8, key


def test_graph_replay_equals_eager(c9k, capi):
    pos, rad, L = c9k
    E = capi.System(pos, rad, L)
    G = capi.System(pos, rad, L)
    for s in (E, G):
        s.first_update()
    E.step(3)
    G.step(3, graph=True)
    a, b = E.state(), G.state()
    for key in ("pos", "vel", "ang_vel", "na_vel", "na_ang_vel", "shear_disp"):
        assert np.array_equal(a[key], b[key]), key
    assert (a["dt"], a["strain"]) == (b["dt"], b["strain"])


def test_large_instance_in_a_batch_equals_its_lone_run(c9k, capi):
    pos, rad, L = c9k
    small = conf(500, 0.52, 7)
    confs = [(pos, rad, L), small]
    slots = cells = 0
    for p, r, l in confs:
        sl, ce, _ = capi.footprint(len(r), r, l)
        slots += sl
        cells += ce
    B = capi.Batch(slots, cells, 2)
    ids = [B.add(p, r, l) for p, r, l in confs]
    B.step(3)
    B.sync()
    assert B.solver_slots(ids[0]).min() >= 0
    for k, (p, r, l) in enumerate(confs):
        S = capi.System(p, r, l)
        S.first_update()
        S.step(3)
        a, b = B.state(ids[k]), S.state()
        for key in ("pos", "vel", "ang_vel", "na_vel", "na_ang_vel", "shear_disp"):
            assert np.array_equal(a[key], b[key]), (k, key)


def test_sixteen_wave_form_on_small_instances(capi, oracle):
    """agg_waves = 16 puts instances of any size on the large-instance form:
    a batch of three equals their lone runs bit for bit, the velocities at
    t=0 are within 1e-8 of the oracle's FP64 direct solve, and under stress
    control both right-hand sides in one launch agree with one per launch."""
    confs = [conf(500, 0.50, 11), conf(1500, 0.56, 12), conf(3000, 0.55, 13)]
    opts = capi.default_opts(precond=1, agg_waves=16)
    slots = cells = 0
    for p, r, l in confs:
        sl, ce, _ = capi.footprint(len(r), r, l)
        slots += sl
        cells += ce
    B = capi.Batch(slots, cells, len(confs), opts=opts)
    ids = [B.add(p, r, l) for p, r, l in confs]
    B.step(3)
    B.sync()
    for k, (p, r, l) in enumerate(confs):
        S = capi.System(p, r, l, opts=opts)
        S.first_update()
        S.step(3)
        assert S.precond_waves() > 0
        a, b = B.state(ids[k]), S.state()
        for key in ("pos", "vel", "ang_vel", "na_vel", "na_ang_vel", "shear_disp"):
            assert np.array_equal(a[key], b[key]), (k, key)
    p, r, l = confs[1]
    S = capi.System(p, r, l, opts=opts)
    R = oracle.System(p, r, l)
    S.first_update()
    R.first_update()
    S.compute_velocities(True)
    R.set_in_predictor(True)
    R.compute_velocities()
    for key in ("na_vel", "na_ang_vel"):
        assert _rel(S.state()[key], R.state()[key]) <= 1e-8, key
    out = []
    for m in (0, 1):
        S = capi.System(p, r, l, opts=capi.default_opts(precond=1, agg_waves=16, max_rhs=m))
        S.first_update()
        if m == 0:
            S.compute_stress(False)
            target = S.stress()["total"][2]
        S.set_control("stress", target)
        S.step(2)
        out.append(S.state())
    for key in ("na_vel", "na_ang_vel"):
        assert np.abs(out[0][key] - out[1][key]).max() <= 1e-9 * np.abs(out[1][key]).max(), key


def test_new_instance_over_a_retired_ones_barrier_words(capi):
    """A batch whose layout changes: P (4 row blocks), Q and R after it; R
    steps (its barrier generation word advances), Q and R retire, and S,
    larger, is placed right after P, where its barrier leaf counters fall on
    R's old words.  S must solve (every barrier word starts from zero when
    the job table is rebuilt) and equal its lone run."""
    opts = capi.default_opts(precond=1, agg_waves=16)
    P, Q, R, S = conf(500, 0.52, 21), conf(100, 0.50, 22), conf(2000, 0.55, 23), conf(5000, 0.55, 24)
    fp = [capi.footprint(len(c[1]), c[1], c[2]) for c in (P, Q, R, S)]
    slots = fp[0][0] + fp[1][0] + fp[2][0] + fp[3][0]
    cells = fp[0][1] + fp[1][1] + fp[2][1] + fp[3][1]
    B = capi.Batch(slots, cells, 4, opts=opts)
    ip, iq, ir = (B.add(*c) for c in (P, Q, R))
    B.step(3)
    B.sync()
    B.remove(iq)
    B.remove(ir)
    i_s = B.add(*S)
    B.step(3)
    B.sync()
    lone = capi.System(*S, opts=opts)
    lone.first_update()
    lone.step(3)
    a, b = B.state(i_s), lone.state()
    for key in ("pos", "vel", "ang_vel", "na_vel", "na_ang_vel", "shear_disp"):
        assert np.array_equal(a[key], b[key]), key
