"""Served projection stream: continuous batching of mixed-size requests.

An ADMM/SDP solver hands the server symmetric float32 matrices that already
live on the device and collects their projections (the reference's intended
caller, include/psd_projection/check.h:10, projects cone blocks every
iteration).  Requests are admitted in arrival order while their memory fits
an HBM budget and run at once:

* small requests (n <= small_max) join a session (psdp_session_*): every
  tick runs ONE filter-chain product of every request in it as one grouped
  launch per operand format, so a request that arrives while others are on
  the GPU starts at the next product boundary (after its own Lanczos bound,
  which runs on a second stream beside the products) instead of waiting for
  the running batch to finish, and leaves after its own output product;
* large requests run one per lane (psdp_project_f32, one handle and stream
  per lane), so a large projection never blocks the small ones.

A request's result does not depend on what else is in flight: a session
block gets bitwise the result psdp_project_batched_f32 gives it alone, and a
lane runs the single-matrix kernels on its own stream.

Memory: the budget covers what the executor holds for the whole run (the
session's arena and the lanes' workspaces, sized once at warm-up) plus the
outputs of admitted requests until the caller releases them (release(), done
automatically after on_done returns).  A small request that does not fit the
session's arena yet waits in the queue; a device allocation that fails at
admission (torch's OutOfMemoryError) requeues the request instead of
crashing the loop.

The executor is pluggable (``TorchExecutor`` on the GPU; tests drive the
scheduler with a fake one on the CPU).
"""
from __future__ import annotations

import collections
import dataclasses
import time
from typing import Any, Callable, Deque, Dict, List, Optional


@dataclasses.dataclass
class Request:
    rid: int
    n: int
    X: Any                      # device tensor (n x n, float32), read only
    t_arrival: float            # host seconds (time.perf_counter clock)
    Y: Any = None
    bytes: int = 0              # output bytes charged to the budget until released
    lane: int = -1              # large lane, or -1 for the session
    t_launch: float = 0.0
    t_done: float = 0.0
    err: Any = None             # optional device scalar (relative error), set by a callback
    released: bool = False
    joined_with: int = -1       # session requests still in flight when this one joined (small requests)

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrival


class TorchExecutor:
    """Session (small requests) and lanes (large requests) on HIP streams
    through libpsdp (GPU)."""

    # One large lane by default.  Measured on the 5-size stream at 390/s: with
    # two lanes side by side the large products hold every CU more of the time
    # and the session's short kernels wait for them (small requests' p50 30 ms,
    # p99 285 ms, burst capacity 643/s); with one lane 9.5 / 45 ms at 594/s.
    # Without large requests in the stream the session alone gives 1.1 / 3.3 ms.
    def __init__(self, device=0, method="mixed", large_lanes=1, seed=5, session_bytes=6 << 30,
                 small_max=1024, ticks_ahead=4, session_method=None):
        import torch
        from . import Handle, Session, default_config, workspace_size
        self.torch = torch
        self.dev = torch.device("cuda", device)
        self.cfg = default_config(method, synchronize=0)  # the lanes' schedule
        self.method = method
        # the session's schedule (small requests), the lanes' unless given
        self.session_method = session_method or method
        self.sess_cfg = default_config(self.session_method, synchronize=0)
        self.small_max = small_max
        self.ticks_ahead = ticks_ahead
        self._ws = workspace_size
        self.sess_stream = torch.cuda.Stream(device=self.dev)
        self.sess_handle = Handle(device, stream=self.sess_stream, seed=seed)
        self.session = Session(self.sess_handle, session_bytes, config=self.sess_cfg)
        self.session_bytes = int(session_bytes)
        self.tickets: Dict[int, Request] = {}
        self.lanes = []
        for _ in range(large_lanes):
            s = torch.cuda.Stream(device=self.dev)
            self.lanes.append((s, Handle(device, stream=s, seed=seed)))
        self.lane_bytes = 0

    # ------------------------------------------------------------------ memory
    def fixed_bytes(self) -> int:
        return self.session_bytes + self.lane_bytes

    def output_bytes(self, n: int) -> int:
        return 4 * n * n

    def fits_session(self, n: int) -> bool:
        return self.session.block_bytes(n) <= self.session_bytes

    def _alloc_out(self, r: Request) -> bool:
        torch = self.torch
        try:
            r.Y = torch.empty((r.n, r.n), dtype=torch.float32, device=self.dev)
            return True
        except torch.OutOfMemoryError:
            r.Y = None
            return False

    # ------------------------------------------------------------------ session
    def session_add(self, r: Request) -> bool:
        """Admits r into the session; False when the arena (or the output
        allocation) has no room now."""
        if not self._alloc_out(r):
            return False
        t = self.session.add(r.X, r.Y)
        if t is None:
            r.Y = None
            return False
        self.tickets[t] = r
        return True

    def session_step(self) -> List[Request]:
        """Returns the requests whose outputs completed; enqueues the next tick
        (and the prologue of new requests) while fewer than ticks_ahead are
        in flight, so that a request added now joins within that many ticks."""
        done, active, inflight = self.session.poll()
        out = [self.tickets.pop(t) for t in done]
        if active and inflight < self.ticks_ahead:
            self.session.step()
        return out

    def session_active(self) -> bool:
        return bool(self.tickets)

    def session_blocks(self) -> int:
        return len(self.tickets)

    # ------------------------------------------------------------------ lanes
    def launch(self, lane: int, r: Request):
        torch = self.torch
        from . import project
        s, h = self.lanes[lane]
        with torch.cuda.stream(s):
            project(h, r.X, out=r.Y, config=self.cfg)
            ev = torch.cuda.Event()
            ev.record(s)
        return ev

    def alloc_large(self, r: Request) -> bool:
        return self._alloc_out(r)

    def done(self, token) -> bool:
        return token.query()

    def stream_of(self, r: Request):
        return self.sess_stream if r.lane < 0 else self.lanes[r.lane][0]

    def warm(self, sizes):
        """Tunes the lanes' tile sizes and sizes their arenas and tile lists for
        the largest requests (keeps both out of the timed loop; their bytes
        are then charged to the budget once)."""
        torch = self.torch
        large = sorted({n for n in sizes if n > self.small_max})
        for lane in range(len(self.lanes)):
            s, h = self.lanes[lane]
            for n in large:
                h.autotune(n, config=self.cfg)
                if n >= 4096 and len(self.lanes) > 1:
                    # the lanes run side by side, so one projection need not
                    # fill the GPU with its own tiles: 256 x 256 tiles (the
                    # fastest per tile) gave 659 against 643 projections/s of
                    # burst capacity on the 5-size stream
                    for f in ("f16x3", "f16", "bf16x3", "bf16"):
                        h.set_tile(n, f, 256)
                x = torch.eye(n, device=self.dev)
                r = Request(-1, n, x, 0.0, Y=torch.empty_like(x))
                torch.cuda.synchronize()
                self.launch(lane, r)
        if large:
            self.lane_bytes = len(self.lanes) * max(self._ws(n, self.method) for n in large)
        small = sorted({n for n in sizes if n <= self.small_max})
        if small:  # one pass of the session over every small size (prologue and grouped kernels warm)
            xs = [torch.eye(n, device=self.dev) for n in small]
            ys = [torch.empty_like(x) for x in xs]
            torch.cuda.synchronize()
            pending = len(xs)
            for x, y in zip(xs, ys):
                assert self.session.add(x, y) is not None
            while pending:
                done, _, _ = self.session.poll()
                pending -= len(done)
                self.session.step()
        torch.cuda.synchronize()


class ProjectionServer:
    """Admission by HBM budget + continuous batching (session + lanes)."""

    def __init__(self, executor, hbm_budget: int, small_max: int = 1024,
                 on_done: Optional[Callable[[Request], None]] = None, clock=time.perf_counter):
        self.ex = executor
        self.budget = int(hbm_budget)
        self.small_max = small_max
        self.on_done = on_done
        self.clock = clock
        self.pending: Deque[Request] = collections.deque()     # arrived, not admitted
        self.admitted_large: Deque[Request] = collections.deque()
        self.inflight = {}                                       # lane -> (token, req)
        self.used = 0                                            # outputs held (admitted, not released)
        self.peak_used = 0
        self.completed: List[Request] = []
        self.batches = 0
        self.refusals = 0
        if self.ex.fixed_bytes() > self.budget:
            raise ValueError(f"the executor holds {self.ex.fixed_bytes()} B, more than the budget {self.budget}")

    # ------------------------------------------------------------------ api
    def submit(self, req: Request):
        req.bytes = self.ex.output_bytes(req.n)
        if self.ex.fixed_bytes() + req.bytes > self.budget:
            raise ValueError(f"request {req.rid} (n={req.n}) needs {req.bytes} B, more than the budget leaves")
        if req.n <= self.small_max and not self.ex.fits_session(req.n):
            raise ValueError(f"request {req.rid} (n={req.n}) does not fit the session arena")
        self.pending.append(req)

    def release(self, req: Request):
        """The caller is done with req.Y: its bytes leave the budget."""
        if not req.released:
            req.released = True
            req.Y = None
            self.used -= req.bytes

    def held_bytes(self) -> int:
        return self.ex.fixed_bytes() + self.used

    def idle(self) -> bool:
        return not (self.pending or self.admitted_large or self.inflight or self.ex.session_active())

    def _finish(self, r: Request, now: float):
        r.t_done = now
        self.completed.append(r)
        if self.on_done is not None:
            self.on_done(r)
            self.release(r)

    def step(self):
        """Retire finished work, admit in arrival order, advance the session,
        launch idle lanes."""
        now = self.clock()
        for lane in list(self.inflight):
            token, r = self.inflight[lane]
            if self.ex.done(token):
                del self.inflight[lane]
                self._finish(r, now)
        # admission: FIFO while the head's output fits the budget (and, for a
        # small request, the session's arena)
        while self.pending:
            r = self.pending[0]
            if self.held_bytes() + r.bytes > self.budget:
                break
            if r.n <= self.small_max:
                r.joined_with = self.ex.session_blocks()
                if not self.ex.session_add(r):
                    self.refusals += 1
                    break
                r.t_launch = self.clock()
            else:
                if not self.ex.alloc_large(r):
                    self.refusals += 1
                    break
                self.admitted_large.append(r)
            self.pending.popleft()
            self.used += r.bytes
            self.peak_used = max(self.peak_used, self.held_bytes())
        # the session: one product of every small request per tick
        for r in self.ex.session_step():
            self._finish(r, self.clock())
        # lanes
        for lane in range(len(self.ex.lanes)):
            if lane in self.inflight or not self.admitted_large:
                continue
            r = self.admitted_large.popleft()
            r.lane = lane
            r.t_launch = self.clock()
            self.inflight[lane] = (self.ex.launch(lane, r), r)
            self.batches += 1

    def run(self, requests: List[Request], arrival_offsets: List[float], poll_sleep: float = 0.0):
        """Replays an arrival schedule (seconds from start) and serves it to
        completion.  Returns the completed requests (with latencies)."""
        order = sorted(range(len(requests)), key=lambda i: arrival_offsets[i])
        t0 = self.clock()
        k = 0
        while k < len(order) or not self.idle():
            now = self.clock()
            while k < len(order) and t0 + arrival_offsets[order[k]] <= now:
                r = requests[order[k]]
                r.t_arrival = t0 + arrival_offsets[order[k]]
                self.submit(r)
                k += 1
            self.step()
            if poll_sleep:
                time.sleep(poll_sleep)
        return self.completed


def latency_stats(done: List[Request]):
    lat = sorted(r.latency for r in done)
    if not lat:
        return {}

    def pct(p):
        i = min(len(lat) - 1, max(0, int(round(p / 100.0 * (len(lat) - 1)))))
        return lat[i]

    t_first = min(r.t_arrival for r in done)
    t_last = max(r.t_done for r in done)
    return {"requests": len(lat), "p50_ms": pct(50) * 1e3, "p99_ms": pct(99) * 1e3, "max_ms": lat[-1] * 1e3,
            "mean_ms": sum(lat) / len(lat) * 1e3, "proj_per_s": len(lat) / (t_last - t_first)}
