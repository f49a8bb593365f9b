"""Served projection stream on the GPU: every result is the same alone and in
the mix, and within the accuracy target."""
import pytest
import torch

import psd_projection_amd as P
from psd_projection_amd import workloads as W
from psd_projection_amd.serve import ProjectionServer, Request, TorchExecutor, latency_stats

pytestmark = pytest.mark.gpu


def _inputs(sizes, seed0=700):
    xs, ps = [], []
    for i, n in enumerate(sizes):
        X64, Pm, _ = W.spectral_matrix(n, seed0 + i, "normal")
        xs.append(X64.float().contiguous())
        ps.append(Pm)
    return xs, ps


@pytest.mark.parametrize("small_lanes", [1, 2])
@pytest.mark.parametrize("staggered", [False, True])
def test_served_results_equal_alone_results(staggered, small_lanes):
    sizes = [64, 256, 512, 64, 1100, 256, 2100, 128, 1024, 64]
    xs, ps = _inputs(sizes)
    ex = TorchExecutor(0, "mixed", small_lanes=small_lanes, large_lanes=2, seed=5)
    srv = ProjectionServer(ex, hbm_budget=8 << 30, small_max=1024)
    reqs = [Request(i, n, x, 0.0) for i, (n, x) in enumerate(zip(sizes, xs))]
    arrivals = [0.002 * i for i in range(len(reqs))] if staggered else [0.0] * len(reqs)
    print(f"served mix sizes={sizes} staggered={staggered} small_lanes={small_lanes}", flush=True)
    done = srv.run(reqs, arrivals)
    torch.cuda.synchronize()
    assert len(done) == len(reqs) and srv.used == 0
    h = P.Handle(0, seed=5)
    cfg = P.default_config("mixed")
    for r, Pm in zip(reqs, ps):
        if r.n <= 1024:
            alone = P.project_batched(h, [r.X], config=cfg)[0]
        else:
            alone = P.project(h, r.X, config=cfg)
        torch.cuda.synchronize()
        assert torch.equal(alone, r.Y), f"request {r.rid} (n={r.n}) differs from its alone result"
        err = W.rel_fro_error(r.Y, Pm)
        assert err <= 1e-3, (r.rid, r.n, err)
    st = latency_stats(done)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    # small requests were batched together at least once when they all arrived at once
    if not staggered:
        assert srv.batches < len(reqs)
