"""Request-loop logic of lf_dem_amd.serve on the CPU, over a stand-in batch
that advances every live instance's strain by a fixed step (the GPU batch
is covered by tests/test_gpu_serve.py): admission in arrival order by
footprint, retirement at the request's strain, the drain after the arrival
window, and a summary whose percentiles of an empty sample are None."""
import numpy as np
import pytest

from lf_dem_amd import serve


class FakeBatch:
    """Slots and strains only; one step adds `dstrain` to every live instance."""

    def __init__(self, slot_cap, cell_cap, inst_cap, opts=None, dstrain=0.01):
        self.cap, self.used, self.live, self.next, self.d = slot_cap, 0, {}, 0, dstrain

    def info(self):
        return {"slot_cap": self.cap}

    def room(self, slots, cells):
        return self.used + slots <= self.cap

    def add(self, pos, radius, L, params=None, strain_end=-1.0):
        iid = self.next
        self.next += 1
        self.live[iid] = [len(radius), 0.0]
        self.used += len(radius)
        return iid

    def remove(self, iid):
        self.used -= self.live.pop(iid)[0]

    def step(self, n=1, graph=False):
        for v in self.live.values():
            v[1] += self.d * n

    def status(self):
        ids = np.array(sorted(self.live), dtype=np.int64)
        return ids, np.array([self.live[i][1] for i in ids]), np.zeros(len(ids), dtype=np.int32)

    def state(self, iid):
        return {"strain": self.live[iid][1]}


def _server(monkeypatch, slot_cap):
    monkeypatch.setattr(serve.capi, "Batch", FakeBatch)
    monkeypatch.setattr(serve.capi, "footprint", lambda n, rad, L, p=None: (len(rad), 1, 0))
    monkeypatch.setattr(serve.capi, "default_params", lambda **kw: None)
    return serve.Server(slot_cap, slot_cap, 16)


def _req(rid, n, strain, arrival):
    return serve.Request(rid, np.zeros((n, 3)), np.ones(n), np.ones(3), strain=strain, arrival=arrival)


@pytest.mark.timeout(60)
def test_admission_order_retirement_and_drain(monkeypatch):
    srv = _server(monkeypatch, 300)
    t = [0.0]

    def clock():                          # every reading advances the fake clock by 1 ms
        t[0] += 1e-3
        return t[0]

    def step_clock(now_fn=None):          # each batch step takes 0.1 s of the fake clock
        t[0] += 0.1
        return serve.Server.step(srv, now_fn=clock)
    srv.step = step_clock
    # 200 + 200 do not fit together: the second waits for the first to retire
    reqs = [_req(0, 200, 0.05, 0.0), _req(1, 200, 0.03, 0.05), _req(2, 50, 0.02, 0.35), _req(3, 50, 0.02, 5.0)]
    elapsed = srv.run(reqs, duration=1.0, drain=10.0, clock=clock)
    s = srv.summary(elapsed)
    assert [r.rid for r in srv.done] == [0, 2, 1]       # 3 arrives after the window and is never served
    assert reqs[1].admitted >= reqs[0].finished - 1e-12
    assert [r.steps for r in (reqs[0], reqs[1], reqs[2])] == [5, 3, 2]
    assert s["completed"] == 3 and s["waiting_at_end"] == 0 and s["live_at_end"] == 0
    assert s["response_s"]["p50"] > 0 and s["steps_per_request"]["p50"] == 3


def test_summary_of_an_empty_run_has_no_response_times(monkeypatch):
    srv = _server(monkeypatch, 300)
    s = srv.summary(1.0)
    assert s["completed"] == 0 and s["completed_per_s"] == 0
    assert s["response_s"] is None and s["admission_wait_s"] is None and s["step_ms"] is None
