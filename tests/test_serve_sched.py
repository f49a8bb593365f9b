"""Scheduler of the served projection stream, driven on the CPU by a fake
executor with a virtual clock (no GPU, no library calls): a session that
advances every block one product per tick (22 products, as the mixed
schedule) and large lanes that finish after a modelled time."""
import random

import pytest

from psd_projection_amd.serve import ProjectionServer, Request, latency_stats

STAGES = 22


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


class FakeExecutor:
    def __init__(self, clock, large_lanes=2, arena=64 << 20, tick=1e-4, lane_bytes=1 << 20, small_max=1024,
                 fail_allocs=0):
        self.clock = clock
        self.lanes = [None] * large_lanes
        self.arena = arena
        self.arena_used = 0
        self.tick = tick
        self.next_tick = 0.0
        self.lane_bytes = lane_bytes
        self.small_max = small_max
        self.blocks = {}          # rid -> [req, products left]
        self.ticks = []           # (time, [rids in the tick])
        self.launches = []
        self.fail_allocs = fail_allocs

    def fixed_bytes(self):
        return self.arena + len(self.lanes) * self.lane_bytes

    def output_bytes(self, n):
        return 4 * n * n

    def block_bytes(self, n):
        return 26 * n * n

    def fits_session(self, n):
        return self.block_bytes(n) <= self.arena

    def session_add(self, r):
        if self.arena_used + self.block_bytes(r.n) > self.arena:
            return False
        self.arena_used += self.block_bytes(r.n)
        r.Y = ("Y", r.rid)
        self.blocks[r.rid] = [r, STAGES]
        return True

    def session_step(self):
        if not self.blocks or self.clock() < self.next_tick:
            return []
        self.next_tick = self.clock() + self.tick
        self.ticks.append((self.clock(), sorted(self.blocks)))
        out = []
        for rid in list(self.blocks):
            self.blocks[rid][1] -= 1
            if self.blocks[rid][1] == 0:
                r = self.blocks.pop(rid)[0]
                self.arena_used -= self.block_bytes(r.n)
                out.append(r)
        return out

    def session_active(self):
        return bool(self.blocks)

    def session_blocks(self):
        return len(self.blocks)

    def alloc_large(self, r):
        if self.fail_allocs:
            self.fail_allocs -= 1
            return False
        r.Y = ("Y", r.rid)
        return True

    def launch(self, lane, r):
        work = r.n ** 3 * 1e-12 + 1e-4
        self.launches.append((lane, r.rid, self.clock()))
        return self.clock() + work

    def done(self, token):
        return self.clock() >= token


def drive(server, clock, reqs, arrivals, dt=1e-5, release_after=None):
    order = sorted(range(len(reqs)), key=lambda i: arrivals[i])
    k = 0
    steps = 0
    while k < len(order) or not server.idle():
        while k < len(order) and arrivals[order[k]] <= clock.t:
            r = reqs[order[k]]
            r.t_arrival = arrivals[order[k]]
            server.submit(r)
            k += 1
        server.step()
        assert server.held_bytes() <= server.budget
        if release_after is not None:
            for r in server.completed:
                if not r.released and clock.t >= r.t_done + release_after:
                    server.release(r)
        clock.t += dt
        steps += 1
        assert steps < 10_000_000
    return server.completed


def make(nreq, seed=5):
    rng = random.Random(seed)
    sizes = [64, 256, 1024, 4096, 8192]
    w = [0.4, 0.3, 0.15, 0.1, 0.05]
    reqs, arr, t = [], [], 0.0
    for i in range(nreq):
        n = rng.choices(sizes, w)[0]
        t += rng.expovariate(2000.0)
        reqs.append(Request(i, n, None, 0.0))
        arr.append(t)
    return reqs, arr


def test_every_request_completes_within_budget_and_routes_by_size():
    clock = Clock()
    ex = FakeExecutor(clock)
    budget = ex.fixed_bytes() + 3 * 4 * 8192 * 8192
    srv = ProjectionServer(ex, budget, small_max=1024, on_done=lambda r: None, clock=clock)
    reqs, arr = make(300)
    done = drive(srv, clock, reqs, arr)
    assert sorted(r.rid for r in done) == list(range(300))
    assert srv.used == 0 and srv.peak_used <= budget
    for lane, rid, _ in ex.launches:
        assert reqs[rid].n > 1024
    small = {r.rid for r in reqs if r.n <= 1024}
    in_ticks = {rid for _, rids in ex.ticks for rid in rids}
    assert small == in_ticks
    # continuous batching: ticks carry requests at different points of their chains
    assert any(len(rids) > 1 for _, rids in ex.ticks)
    for r in done:
        assert r.t_done >= r.t_launch >= r.t_arrival
    st = latency_stats(done)
    assert st["requests"] == 300 and st["p99_ms"] >= st["p50_ms"] > 0


def test_late_request_joins_at_the_next_product():
    # A arrives alone; B arrives while A is in the middle of its 22 products.
    # B joins at the next tick (it does not wait for A to finish) and
    # completes 22 ticks after it joined: as many ticks after A as it joined
    # after A started.
    clock = Clock()
    ex = FakeExecutor(clock, tick=1e-3)
    srv = ProjectionServer(ex, ex.fixed_bytes() + (1 << 20), on_done=lambda r: None, clock=clock)
    a, b = Request(0, 64, None, 0.0), Request(1, 64, None, 0.0)
    done = drive(srv, clock, [a, b], [0.0, 10.5e-3])
    first_b = next(t for t, rids in ex.ticks if 1 in rids)
    ticks_a = [t for t, rids in ex.ticks if 0 in rids]
    assert len(ticks_a) == STAGES
    assert ticks_a[0] < first_b < ticks_a[-1]            # B ran beside A
    assert first_b - 10.5e-3 <= 1.01e-3                  # within one tick of its arrival
    assert b.t_done - a.t_done == pytest.approx(first_b - ticks_a[0], abs=5e-5)
    assert first_b - ticks_a[0] < 12e-3
    assert len(done) == 2
    assert (a.joined_with, b.joined_with) == (0, 1)      # B was admitted with A in flight


def test_outputs_count_until_released_and_admission_is_fifo():
    # Room for the fixed memory plus two n = 1024 outputs: the third request
    # waits until the caller releases a finished one (results that are still
    # held keep their memory), and admission stays in arrival order.
    clock = Clock()
    ex = FakeExecutor(clock)
    out = 4 * 1024 * 1024
    srv = ProjectionServer(ex, ex.fixed_bytes() + 2 * out, clock=clock)   # no on_done: caller releases
    reqs = [Request(i, 1024, None, 0.0) for i in range(3)] + [Request(3, 64, None, 0.0)]
    done = drive(srv, clock, reqs, [0.0] * 4, release_after=5e-3)
    assert [r.rid for r in sorted(done, key=lambda r: r.t_launch)] == [0, 1, 2, 3]
    assert reqs[2].t_launch >= reqs[0].t_done + 5e-3 - 1e-9 or reqs[2].t_launch >= reqs[1].t_done + 5e-3 - 1e-9
    assert srv.peak_used <= srv.budget
    for r in done:
        srv.release(r)
    assert srv.used == 0


def test_allocation_failure_requeues():
    clock = Clock()
    ex = FakeExecutor(clock, fail_allocs=3)
    srv = ProjectionServer(ex, ex.fixed_bytes() + (1 << 30), on_done=lambda r: None, clock=clock)
    done = drive(srv, clock, [Request(0, 4096, None, 0.0)], [0.0])
    assert len(done) == 1 and srv.refusals == 3


def test_session_arena_full_waits():
    clock = Clock()
    ex = FakeExecutor(clock, arena=2 * 26 * 1024 * 1024)
    srv = ProjectionServer(ex, ex.fixed_bytes() + (1 << 30), on_done=lambda r: None, clock=clock)
    reqs = [Request(i, 1024, None, 0.0) for i in range(4)]
    done = drive(srv, clock, reqs, [0.0] * 4)
    assert len(done) == 4
    assert max(len(rids) for _, rids in ex.ticks) == 2     # never more than the arena holds


def test_oversized_request_rejected():
    clock = Clock()
    ex = FakeExecutor(clock)
    srv = ProjectionServer(ex, ex.fixed_bytes() + (1 << 20), clock=clock)
    with pytest.raises(ValueError):
        srv.submit(Request(0, 4096, None, 0.0))
    with pytest.raises(ValueError):
        ProjectionServer(ex, ex.fixed_bytes() - 1, clock=clock)
