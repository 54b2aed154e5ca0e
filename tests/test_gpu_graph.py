"""Captured steps (lfd_system_step_graph, lfd_batch_step_graph) against eager
steps through the C ABI.

A captured step replays the same kernels with the same arguments, so the bar
is bit-identical states.  Every replay is checked: the graphs of both
pair-buffer parities are replayed several times, with the program's own eager
work (state and pair reads, synchronisations) between the replays, and again
after an option change that forces a new capture.
"""
import numpy as np
import pytest

from tests.conftest import conf

pytestmark = pytest.mark.gpu


def _same(a, b):
    for k in ("pos", "vel", "ang_vel", "na_vel", "na_ang_vel", "shear_disp"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("dt", "time", "strain"):
        assert a[k] == b[k], k


@pytest.mark.parametrize("precond", [1, 0])
def test_graph_steps_match_eager(precond, capi):
    pos, rad, L = conf(500, 0.5, 1)
    opts = capi.default_opts(precond=precond)
    E = capi.System(pos, rad, L, opts=opts)
    G = capi.System(pos, rad, L, opts=opts)
    E.first_update()
    G.first_update()
    for k in range(7):
        print("preconditioner", precond, "step", k, flush=True)
        E.step(1)
        G.step(1, graph=True)
        _same(E.state(), G.state())
        pe, pg = E.pairs(), G.pairs()
        assert np.array_equal(pe["p0"], pg["p0"]) and np.array_equal(pe["p1"], pg["p1"])
    assert G.graphs() == 2 and E.graphs() == 0
    E.step(3)
    G.step(3, graph=True)
    _same(E.state(), G.state())
    # new options: the job table changes and the step is captured again
    o2 = capi.default_opts(precond=precond, regroup_interval=2, inner_tol=3e-3)
    E.set_opts(o2)
    G.set_opts(o2)
    for k in range(3):
        print("after option change, step", k, flush=True)
        E.step(1)
        G.step(1, graph=True)
        _same(E.state(), G.state())


def test_graph_falls_back_for_the_plain_solver(capi):
    """The plain FP64 form has host round trips inside a solve: its graph
    steps run eagerly and give the eager result."""
    pos, rad, L = conf(300, 0.5, 2)
    opts = capi.default_opts(kind=1, precond=0)
    E = capi.System(pos, rad, L, opts=opts)
    G = capi.System(pos, rad, L, opts=opts)
    E.first_update()
    G.first_update()
    E.step(2)
    G.step(2, graph=True)
    _same(E.state(), G.state())
    assert G.graphs() == 0


def test_batch_graph_steps_match_eager(capi):
    specs = [(300, 0.50, 11, 5.0), (700, 0.55, 12, 0.5)]
    batches = []
    for _ in range(2):
        slots = cells = 0
        confs = []
        for n, phi, seed, rep in specs:
            pos, rad, L = conf(n, phi, seed)
            p = capi.default_params(repulsion=rep)
            sl, ce, _ = capi.footprint(len(rad), rad, L, p)
            slots += sl
            cells += ce
            confs.append((pos, rad, L, p))
        B = capi.Batch(slots, cells, len(specs))
        ids = [B.add(pos, rad, L, params=p) for pos, rad, L, p in confs]
        batches.append((B, ids))
    (E, ie), (G, ig) = batches
    for k in range(5):
        print("batch step", k, flush=True)
        E.step(1)
        G.step(1, graph=True)
        E.sync()
        G.sync()
        for a, b in zip(ie, ig):
            _same(E.state(a), G.state(b))
    assert G.graphs() == 2
