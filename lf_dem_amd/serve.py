"""Continuous-batching request loop over one lfd_batch.

Simulation requests (a configuration, parameters and a strain to run) arrive
at any time.  Between two batch steps the loop admits waiting requests, in
arrival order, while the batch has room for their slots and cells (their HBM
footprint, lfd_batch_footprint); it retires every instance that has reached
its strain and hands back its final state.  Admitted requests are stepped
together with the ones already running: positions, contact states and the
preconditioner data stay resident on the device for the request's whole life,
and every batch step solves all live instances in the same cooperative
launches.  A request's trajectory is the one it has alone (instances share no
arithmetic), so the result does not depend on what else was in the batch.

LF_DEM has no counterpart: it runs one System per process
(Simulation.cpp:241-264 drives a single System).
"""
import collections
import time

import numpy as np

from . import capi


class Request:
    """One simulation: pos (n, 3), radius (n,), box L, parameters, strain to run."""

    __slots__ = ("rid", "pos", "radius", "L", "params", "strain", "arrival", "admitted", "finished",
                 "steps", "status", "state", "slots", "cells", "loader")

    def __init__(self, rid, pos, radius, L, params=None, strain=0.5, arrival=0.0, loader=None):
        """pos may be None with `loader` a callable returning the positions,
        called once at admission (the footprint needs only radius and L)."""
        self.rid = rid
        self.loader = loader
        self.pos = None if pos is None else np.ascontiguousarray(pos, dtype=np.float64)
        self.radius = np.ascontiguousarray(radius, dtype=np.float64)
        self.L = np.ascontiguousarray(L, dtype=np.float64)
        self.params = params if params is not None else capi.default_params()
        self.strain = float(strain)
        self.arrival = float(arrival)
        self.admitted = None
        self.finished = None
        self.steps = 0
        self.status = None
        self.state = None
        self.slots, self.cells, _ = capi.footprint(len(self.radius), self.radius, self.L, self.params)


class Server:
    """Request loop over a batch of fixed capacity (slots, cells, handles)."""

    def __init__(self, slot_cap, cell_cap, inst_cap, opts=None, keep_state=True, graph=False):
        """graph: step by replaying captured step graphs (lfd_batch_step_graph),
        recaptured whenever the batch's membership changes."""
        self.batch = capi.Batch(slot_cap, cell_cap, inst_cap, opts=opts)
        self.graph = graph
        self.waiting = collections.deque()
        self.live = {}                 # instance id -> Request
        self.done = []
        self.step_ms = []
        self.keep_state = keep_state
        self.particle_steps = 0
        self.instance_steps = 0

    def submit(self, req):
        if req.slots > self.batch.info()["slot_cap"]:
            raise ValueError("request %s needs more slots than the batch holds" % req.rid)
        self.waiting.append(req)

    def admit(self, now):
        """Admit waiting requests in arrival order while they fit."""
        n = 0
        while self.waiting:
            req = self.waiting[0]
            if not self.batch.room(req.slots, req.cells):
                break
            if req.pos is None:
                req.pos = np.ascontiguousarray(req.loader(), dtype=np.float64)
            iid = self.batch.add(req.pos, req.radius, req.L, params=req.params, strain_end=req.strain)
            req.admitted = now
            self.live[iid] = req
            self.waiting.popleft()
            n += 1
        return n

    def step(self, now_fn=time.perf_counter):
        """One batch step of every live instance, then retirement of the
        finished ones.  Returns the step's wall time in ms."""
        if not self.live:
            return 0.0
        t0 = now_fn()
        self.batch.step(1, graph=self.graph)
        ids, strain, status = self.batch.status()
        ms = (now_fn() - t0) * 1e3
        self.step_ms.append(ms)
        for iid in ids:
            self.live[int(iid)].steps += 1
            self.particle_steps += len(self.live[int(iid)].radius)
        self.instance_steps += len(ids)
        t_end = now_fn()
        for iid, s, st in zip(ids.tolist(), strain.tolist(), status.tolist()):
            req = self.live[iid]
            # System::keepRunning (System.cpp:1189-1201)
            if st != 0 or s > req.strain - 1e-8:
                req.status = st
                req.finished = t_end
                if self.keep_state:
                    req.state = self.batch.state(iid)
                self.batch.remove(iid)
                del self.live[iid]
                self.done.append(req)
        return ms

    def run(self, requests, duration=None, drain=0.0, clock=time.perf_counter):
        """Feeds `requests` (sorted by arrival, seconds from the start) through
        the loop until all are done or `duration` seconds have passed; with
        `drain` > 0 the requests that arrived by then are served for up to
        `drain` seconds more (no new arrivals)."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        t_start = clock()
        while pending or self.waiting or self.live:
            now = clock() - t_start
            if duration is not None and now > duration:
                while pending and pending[0].arrival <= duration:
                    self.submit(pending.popleft())
                pending.clear()
                if now > duration + drain or not (self.waiting or self.live):
                    break
            while pending and pending[0].arrival <= now:
                self.submit(pending.popleft())
            self.admit(now)
            if self.live:
                self.step(lambda: clock() - t_start)
            elif pending:
                time.sleep(min(0.001, max(0.0, pending[0].arrival - now)))
        return clock() - t_start

    def summary(self, elapsed):
        """Step latency, throughput and, over the completed requests, response
        time (arrival to retirement) and admission wait; percentiles of an
        empty sample are None."""
        def pct(x):
            if len(x) == 0:
                return None
            return {"p50": float(np.percentile(x, 50)), "p99": float(np.percentile(x, 99))}
        fin = [r for r in self.done if r.status == 0]
        adm = self.done+list(self.live.values())
        return {
            "batch_steps": len(self.step_ms),
            "step_ms": pct(self.step_ms),
            "instance_steps_per_s": self.instance_steps / elapsed,
            "particle_steps_per_s": self.particle_steps / elapsed,
            "completed": len(fin), "failed": len(self.done) - len(fin),
            "completed_per_s": len(fin) / elapsed,
            "live_at_end": len(self.live), "waiting_at_end": len(self.waiting),
            "response_s": pct([r.finished - r.arrival for r in fin]),
            "steps_per_request": pct([r.steps for r in fin]),
            "admission_wait_s": pct([r.admitted - r.arrival for r in adm]),
        }
