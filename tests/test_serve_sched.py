"""Scheduler of the served projection stream, driven on the CPU by a fake
executor with a virtual clock (no GPU, no library calls)."""
import random

import pytest

from psd_projection_amd.serve import ProjectionServer, Request, latency_stats


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


class FakeExecutor:
    """Lanes finish a launch after a modelled time; records what ran where."""

    def __init__(self, clock, small_lanes=2, large_lanes=2):
        self.clock = clock
        self.small = small_lanes
        self.lanes = [None] * (small_lanes + large_lanes)
        self.launches = []

    def is_small_lane(self, lane):
        return lane < self.small

    def request_bytes(self, n, small):
        return 4 * n * n + (24 if small else 28) * n * n

    def launch(self, lane, reqs):
        work = sum(r.n ** 3 for r in reqs) * 1e-12 + 1e-4
        self.launches.append((lane, [r.rid for r in reqs], self.clock()))
        for r in reqs:
            r.Y = ("Y", r.rid)
        return self.clock() + work

    def done(self, token):
        return self.clock() >= token


def drive(server, clock, reqs, arrivals, dt=1e-4):
    order = sorted(range(len(reqs)), key=lambda i: arrivals[i])
    k = 0
    while k < len(order) or not server.idle():
        while k < len(order) and arrivals[order[k]] <= clock.t:
            r = reqs[order[k]]
            r.t_arrival = arrivals[order[k]]
            server.submit(r)
            k += 1
        server.step()
        assert server.used <= server.budget
        clock.t += dt
    return server.completed


def make(nreq, seed=5):
    rng = random.Random(seed)
    sizes = [64, 256, 1024, 4096, 8192]
    w = [0.4, 0.3, 0.15, 0.1, 0.05]
    reqs, arr, t = [], [], 0.0
    for i in range(nreq):
        n = rng.choices(sizes, w)[0]
        t += rng.expovariate(200.0)
        reqs.append(Request(i, n, None, 0.0))
        arr.append(t)
    return reqs, arr


def test_every_request_completes_within_budget_and_lanes_route_by_size():
    clock = Clock()
    ex = FakeExecutor(clock)
    budget = 3 * (28 * 8192 * 8192 + 4 * 8192 * 8192)
    srv = ProjectionServer(ex, budget, small_max=1024, clock=clock)
    reqs, arr = make(400)
    done = drive(srv, clock, reqs, arr)
    assert sorted(r.rid for r in done) == list(range(400))
    assert srv.used == 0 and srv.peak_used <= budget
    for lane, rids, _ in ex.launches:
        ns = {reqs[i].n for i in rids}
        if ex.is_small_lane(lane):
            assert max(ns) <= 1024
        else:
            assert len(rids) == 1 and min(ns) > 1024
    # continuous batching: small requests that queued during a launch form the next batch
    assert any(len(rids) > 1 for lane, rids, _ in ex.launches if ex.is_small_lane(lane))
    for r in done:
        assert r.t_done >= r.t_launch >= r.t_arrival
    st = latency_stats(done)
    assert st["requests"] == 400 and st["p99_ms"] >= st["p50_ms"] > 0


def test_admission_is_fifo_and_waits_for_memory():
    clock = Clock()
    ex = FakeExecutor(clock, small_lanes=1, large_lanes=1)
    one_big = 28 * 4096 * 4096 + 4 * 4096 * 4096
    srv = ProjectionServer(ex, one_big, small_max=1024, clock=clock)   # room for one n=4096 at a time
    reqs = [Request(i, 4096, None, 0.0) for i in range(3)] + [Request(3, 64, None, 0.0)]
    done = drive(srv, clock, reqs, [0.0, 0.0, 0.0, 0.0])
    launches = [rids for _, rids, _ in ex.launches]
    assert launches[:3] == [[0], [1], [2]]          # one at a time: memory frees between them
    assert launches[3] == [3]                        # the small one waited its turn (FIFO admission)
    assert srv.peak_used <= one_big and len(done) == 4


def test_oversized_request_rejected():
    clock = Clock()
    srv = ProjectionServer(FakeExecutor(clock), 1 << 20, clock=clock)
    with pytest.raises(ValueError):
        srv.submit(Request(0, 4096, None, 0.0))
