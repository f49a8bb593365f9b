"""Served projection stream on the GPU: small requests join the running
session between products, large ones run on lanes; every result is the same
alone and in the mix, and within the accuracy target."""
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
        del X64
    torch.cuda.synchronize()
    return xs, ps


@pytest.mark.parametrize("arrivals", ["burst", "staggered", "mid_batch"])
def test_served_results_equal_alone_results(arrivals):
    sizes = [64, 256, 512, 64, 4096, 1100, 256, 8192, 2100, 128, 1024, 64, 33, 700]
    xs, ps = _inputs(sizes)
    ex = TorchExecutor(0, "mixed", large_lanes=2, seed=5, session_bytes=2 << 30)
    ex.warm(sorted(set(sizes)))
    srv = ProjectionServer(ex, hbm_budget=16 << 30, small_max=1024)
    reqs = [Request(i, n, x, 0.0) for i, (n, x) in enumerate(zip(sizes, xs))]
    if arrivals == "burst":
        offs = [0.0] * len(reqs)
    elif arrivals == "staggered":
        offs = [0.002 * i for i in range(len(reqs))]
    else:  # the last half arrives while the first half is mid-chain (the session runs a product every ~0.1 ms)
        offs = [0.0] * 7 + [0.0015 + 0.0003 * i for i in range(len(reqs) - 7)]
    print(f"served mix sizes={sizes} arrivals={arrivals}", flush=True)
    done = srv.run(reqs, offs)
    torch.cuda.synchronize()
    assert len(done) == len(reqs)
    h = P.Handle(0, seed=5)
    cfg = P.default_config("mixed")
    for r, Pm in zip(reqs, ps):
        if r.n <= 1024:
            alone = P.project_batched(h, [r.X], config=cfg)[0]
        else:  # alone on the lane's own handle (same seed, and the tile sizes its autotuner chose)
            alone = P.project(ex.lanes[r.lane][1], r.X, config=cfg)
        torch.cuda.synchronize()
        assert torch.equal(alone, r.Y), f"request {r.rid} (n={r.n}) differs from its alone result"
        err = W.rel_fro_error(r.Y, Pm)
        assert err <= 1e-3, (r.rid, r.n, err)
        srv.release(r)
    assert srv.used == 0
    st = latency_stats(done)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    if arrivals == "mid_batch":
        # the small requests that arrived mid-chain joined the running session:
        # at least one was admitted while other session requests were still on
        # the GPU (counted by the server at admission, not by wall clock)
        late = [r for r in reqs[7:] if r.n <= 1024]
        assert max(r.joined_with for r in late) > 0, [(r.rid, r.n, r.joined_with) for r in late]
