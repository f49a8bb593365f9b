"""Autotuner and graph replay on the GPU: a tuned or replayed projection is
bitwise the projection of the default path (every tile size accumulates each
entry over k in the same order; a graph replays the same launches)."""
import numpy as np
import pytest
import torch

import psd_projection_amd as P
from psd_projection_amd import workloads as W

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def blocks(sizes, seed0):
    xs, ps = [], []
    for i, n in enumerate(sizes):
        X64, Pm, _ = W.spectral_matrix(n, seed0 + i, "uniform")
        xs.append(X64.float().contiguous())
        ps.append(Pm)
    return xs, ps


@pytest.mark.parametrize("method,n", [("mixed", 1500), ("mixed", 2304),
                                      ("mixed_mx", 2304)])
def test_autotuned_projection_equals_default(method, n):
    X64, Pm, _ = W.spectral_matrix(n, 31, "normal")
    X = X64.float().contiguous()
    h = P.Handle(0, seed=3)
    cfg = P.default_config(method)
    print(f"autotune {method} n={n}", flush=True)
    Y0 = P.project(h, X, config=cfg)
    torch.cuda.synchronize()
    choice = h.autotune(n, config=cfg)
    print(f"  tiles chosen {choice}", flush=True)
    assert choice and all(t in (64, 128, 256) for t in choice.values())
    Y1 = P.project(h, X, config=cfg)
    torch.cuda.synchronize()
    assert torch.equal(Y0, Y1)
    assert W.rel_fro_error(Y1, Pm) <= 1e-3
    h.close()


@pytest.mark.parametrize("method", ["mixed", "mixed_mx"])
def test_batched_autotune_and_tilings_agree(method):
    sizes = [40, 100, 300, 520, 700, 64, 900, 1000, 256, 33]
    xs, ps = blocks(sizes, 900)
    h = P.Handle(0, seed=4)
    print(f"batched autotune {method}", flush=True)
    ref = P.project_batched(h, xs, method=method)
    torch.cuda.synchronize()
    outs = [torch.empty_like(x) for x in xs]
    h.autotune_batched(xs, outs, method=method)
    torch.cuda.synchronize()
    for a, b in zip(ref, outs):
        assert torch.equal(a, b)
    for big_n, conc in ((256, True), (512, False), (100000, False)):
        h.set_batched_tiling(big_n, conc)
        o = P.project_batched(h, xs, method=method)
        torch.cuda.synchronize()
        for a, b in zip(ref, o):
            assert torch.equal(a, b), (big_n, conc)
    for y, pm in zip(ref, ps):
        assert W.rel_fro_error(y, pm) <= 1e-3
    h.close()


@pytest.mark.parametrize("conc", [False, True])
def test_batched_graph_replays_match_eager(conc):
    sizes = [50, 200, 600, 130, 1000, 64]
    xs, ps = blocks(sizes, 950)
    h = P.Handle(0, seed=6)
    h.set_batched_tiling(512, conc)
    eager = P.project_batched(h, xs)
    torch.cuda.synchronize()
    h.set_batched_graph(True)
    outs = [torch.empty_like(x) for x in xs]
    for rep in range(4):
        print(f"graph replay {rep} conc={conc}", flush=True)
        for o in outs:
            o.fill_(float("nan"))
        P.project_batched(h, xs, outs=outs)
        torch.cuda.synchronize()
        for a, b in zip(eager, outs):
            assert torch.equal(a, b), rep
        # eager work between replays
        _ = (xs[0] @ xs[0]).sum().item()
    # new data, same shapes: the replay reads the new descriptors
    xs2, ps2 = blocks(sizes, 990)
    outs2 = P.project_batched(h, xs2)
    torch.cuda.synchronize()
    h.set_batched_graph(False)
    eager2 = P.project_batched(h, xs2)
    torch.cuda.synchronize()
    for a, b in zip(eager2, outs2):
        assert torch.equal(a, b)
    # a different batch shape is captured anew
    h.set_batched_graph(True)
    o3 = P.project_batched(h, xs[:3])
    torch.cuda.synchronize()
    for a, b in zip(eager[:3], o3):
        assert torch.equal(a, b)
    h.close()
