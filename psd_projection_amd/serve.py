"""Served projection stream: continuous batching of mixed-size requests.

An ADMM/SDP solver hands the server symmetric float32 matrices that already
live on the device and collects their projections.  Requests are admitted in
arrival order while their memory fits an HBM budget; admitted requests run on
several lanes (HIP streams) at once, so work that arrives while other work is
on the GPU starts as soon as a lane and memory free up:

* small lanes (n <= small_max): every launch takes all admitted small
  requests at once and projects them with one grouped launch per filter-chain
  product (psdp_project_batched_f32), so requests that queued while the
  previous batch ran form the next batch;
* large lanes: one request per launch (psdp_project_f32), one handle and
  stream per lane, so a large projection never blocks the small ones.

A request's result does not depend on what else is in flight: the grouped
path gives a block bitwise the same result alone and in any batch, and the
single path runs the same kernels on its own stream.

The executor is pluggable (``TorchExecutor`` on the GPU; tests drive the
scheduler with a fake one on the CPU).
"""
from __future__ import annotations

import collections
import dataclasses
import time
from typing import Any, Callable, Deque, List, Optional


@dataclasses.dataclass
class Request:
    rid: int
    n: int
    X: Any                      # device tensor (n x n, float32), read only
    t_arrival: float            # host seconds (time.perf_counter clock)
    Y: Any = None
    bytes: int = 0
    lane: int = -1
    t_launch: float = 0.0
    t_done: float = 0.0
    err: Any = None             # optional device scalar (relative error), set by a callback

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrival


class TorchExecutor:
    """Runs lanes on HIP streams through libpsdp (GPU)."""

    # One small lane: whatever arrives while a batch runs forms the next batch.
    # Measured on the 10,000-request stream at 0.7 x capacity: 1 small + 2
    # large lanes gave p50 9.2 ms / p99 110 ms (small requests' p99 34 ms),
    # 2 + 2 gave p50 20 ms / p99 109 ms (small requests' p99 77 ms) at 3.5%
    # more capacity.
    def __init__(self, device=0, method="mixed", small_lanes=1, large_lanes=2, seed=5):
        import torch
        from . import Handle, default_config, workspace_size, workspace_size_batched
        self.torch = torch
        self.dev = torch.device("cuda", device)
        self.cfg = default_config(method, synchronize=0)
        self.method = method
        self._ws = workspace_size
        self._wsb = workspace_size_batched
        self.small = small_lanes
        self.lanes = []
        for i in range(small_lanes + large_lanes):
            s = torch.cuda.Stream(device=self.dev)
            h = Handle(device, stream=s, seed=seed)
            self.lanes.append((s, h))

    def is_small_lane(self, lane: int) -> bool:
        return lane < self.small

    def request_bytes(self, n: int, small: bool) -> int:
        io = 4 * n * n
        ws = self._wsb([n], self.method) if small else self._ws(n, self.method)
        return io + ws

    def launch(self, lane: int, reqs: List[Request]):
        from . import project, project_batched
        torch = self.torch
        s, h = self.lanes[lane]
        with torch.cuda.stream(s):
            for r in reqs:
                r.Y = torch.empty((r.n, r.n), dtype=torch.float32, device=self.dev)
            if self.is_small_lane(lane):
                project_batched(h, [r.X for r in reqs], outs=[r.Y for r in reqs], config=self.cfg)
            else:
                for r in reqs:
                    project(h, r.X, out=r.Y, config=self.cfg)
            ev = torch.cuda.Event()
            ev.record(s)
        return ev

    def done(self, token) -> bool:
        return token.query()

    def stream_of(self, lane: int):
        return self.lanes[lane][0]

    def warm(self, sizes):
        """Tunes the large lanes' tile sizes and sizes the lane arenas and tile
        lists for the largest requests (keeps both out of the timed loop)."""
        torch = self.torch
        small = [n for n in sizes if n <= 1024]
        large = [n for n in sizes if n > 1024]
        for lane in range(len(self.lanes)):
            if not self.is_small_lane(lane):
                for n in large:  # tile sizes of the single-matrix projections
                    self.lanes[lane][1].autotune(n, config=self.cfg)
        for lane in range(len(self.lanes)):
            ns = small if self.is_small_lane(lane) else large
            if not ns:
                continue
            xs = [torch.eye(n, device=self.dev) for n in ns]
            reqs = [Request(-1, n, x, 0.0) for n, x in zip(ns, xs)]
            if self.is_small_lane(lane):
                self.done(self.launch(lane, reqs))
            else:
                for r in reqs:
                    self.launch(lane, [r])
        torch.cuda.synchronize()


class ProjectionServer:
    """Admission by HBM budget + continuous batching over lanes."""

    def __init__(self, executor, hbm_budget: int, small_max: int = 1024, max_batch: int = 512,
                 on_done: Optional[Callable[[Request], None]] = None, clock=time.perf_counter):
        self.ex = executor
        self.budget = int(hbm_budget)
        self.small_max = small_max
        self.max_batch = max_batch
        self.on_done = on_done
        self.clock = clock
        self.pending: Deque[Request] = collections.deque()     # arrived, not admitted
        self.admitted_small: Deque[Request] = collections.deque()
        self.admitted_large: Deque[Request] = collections.deque()
        self.inflight = {}                                       # lane -> (token, [reqs])
        self.used = 0
        self.peak_used = 0
        self.completed: List[Request] = []
        self.batches = 0

    # ------------------------------------------------------------------ api
    def submit(self, req: Request):
        small = req.n <= self.small_max
        req.bytes = self.ex.request_bytes(req.n, small)
        if req.bytes > self.budget:
            raise ValueError(f"request {req.rid} (n={req.n}) needs {req.bytes} B, more than the budget")
        self.pending.append(req)

    def idle(self) -> bool:
        return not (self.pending or self.admitted_small or self.admitted_large or self.inflight)

    def step(self):
        """Retire finished lanes, admit in arrival order, launch idle lanes."""
        now = self.clock()
        for lane in list(self.inflight):
            token, reqs = self.inflight[lane]
            if self.ex.done(token):
                del self.inflight[lane]
                for r in reqs:
                    r.t_done = now
                    self.used -= r.bytes
                    if self.on_done is not None:
                        self.on_done(r)
                    self.completed.append(r)
        # admission: FIFO while the head fits the budget
        while self.pending and self.used + self.pending[0].bytes <= self.budget:
            r = self.pending.popleft()
            self.used += r.bytes
            self.peak_used = max(self.peak_used, self.used)
            (self.admitted_small if r.n <= self.small_max else self.admitted_large).append(r)
        # launch
        nlanes = len(self.ex.lanes)
        for lane in range(nlanes):
            if lane in self.inflight:
                continue
            if self.ex.is_small_lane(lane):
                if not self.admitted_small:
                    continue
                batch = []
                while self.admitted_small and len(batch) < self.max_batch:
                    batch.append(self.admitted_small.popleft())
            else:
                if not self.admitted_large:
                    continue
                batch = [self.admitted_large.popleft()]
            t = self.clock()
            for r in batch:
                r.lane = lane
                r.t_launch = t
            self.inflight[lane] = (self.ex.launch(lane, batch), batch)
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
