"""Continuous-batching request loop (lf_dem_amd.serve) on the GPU.

Requests arrive while others are being stepped, wait when the batch has no
room for their footprint, and are admitted as earlier ones retire.  Each
request's final state must equal the state of the same simulation run alone
(tolerance 0: instances share no arithmetic), and its step count the lone
run's.
"""
import numpy as np
import pytest

from tests.conftest import conf

pytestmark = pytest.mark.gpu

STRAIN = 0.004


def _req(capi, serve, rid, n, phi, seed, rep, arrival=0.0):
    pos, rad, L = conf(n, phi, seed)
    return serve.Request(rid, pos, rad, L, params=capi.default_params(repulsion=rep), strain=STRAIN,
                         arrival=arrival)


def _alone(capi, req):
    S = capi.System(req.pos, req.radius, req.L, params=req.params)
    S.first_update()
    steps = 0
    while True:
        S.step(1, req.strain)
        steps += 1
        st = S.state()
        if st["strain"] > req.strain - 1e-8:
            return st, steps


def _same(a, b):
    for k in ("pos", "vel", "ang_vel", "na_vel", "na_ang_vel", "shear_disp"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("dt", "time", "strain"):
        assert a[k] == b[k], k


@pytest.mark.parametrize("graph", [False, True])
def test_requests_same_alone_and_in_the_mix(capi, graph):
    """Eager batch steps, and steps replayed from step graphs recaptured at
    every change of the batch's membership."""
    from lf_dem_amd import serve
    reqs = [_req(capi, serve, 0, 300, 0.50, 21, 5.0), _req(capi, serve, 1, 520, 0.55, 22, 0.5),
            _req(capi, serve, 2, 260, 0.45, 23, 50.0), _req(capi, serve, 3, 400, 0.52, 24, 1.0)]
    # room for requests 0 and 1 only: 2 and 3 wait for slots to free up
    slots = reqs[0].slots + reqs[1].slots
    cells = reqs[0].cells + reqs[1].cells
    srv = serve.Server(slots, cells, 4, graph=graph)
    print("graph stepping", graph, flush=True)
    srv.submit(reqs[0])
    srv.admit(0)
    srv.step()
    srv.submit(reqs[1])              # arrives while request 0 is running
    srv.submit(reqs[2])
    srv.submit(reqs[3])
    srv.admit(1)
    assert len(srv.live) == 2 and len(srv.waiting) == 2
    k = 2
    while srv.live or srv.waiting:
        srv.admit(k)
        srv.step()
        k += 1
        assert k < 500
    assert len(srv.done) == 4
    assert all(r.status == 0 for r in srv.done)
    assert reqs[2].admitted > reqs[1].admitted
    for r in reqs:
        st, steps = _alone(capi, r)
        print("request", r.rid, "steps", r.steps, "alone", steps, flush=True)
        assert r.steps == steps
        _same(r.state, st)


def test_footprint_admission(capi):
    from lf_dem_amd import serve
    r = _req(capi, serve, 0, 300, 0.50, 21, 5.0)
    srv = serve.Server(r.slots, r.cells, 2)
    big = _req(capi, serve, 1, 520, 0.55, 22, 0.5)
    with pytest.raises(ValueError):
        srv.submit(big)
    srv.submit(r)
    assert srv.admit(0) == 1
    r2 = _req(capi, serve, 2, 300, 0.50, 25, 5.0)
    srv.submit(r2)
    assert srv.admit(0) == 0          # no slots until the first retires
