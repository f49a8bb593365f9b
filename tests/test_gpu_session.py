"""Continuous batching (psdp_session_*) on the GPU: blocks that join a
running session between products, leave after their output product, and get
bitwise the result psdp_project_batched_f32 gives them alone."""
import time

import pytest

torch = pytest.importorskip("torch")
import psd_projection_amd as P  # noqa: E402
from psd_projection_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _inputs(sizes, seed0=900):
    xs, ps = [], []
    for i, n in enumerate(sizes):
        X64, Pm, _ = W.spectral_matrix(n, seed0 + i, "normal")
        xs.append(X64.float().contiguous())
        ps.append(Pm)
    torch.cuda.synchronize()  # the session runs on its handle's own stream
    return xs, ps


def _drive(sess, waves, max_steps=20000, out=None):
    """Adds each wave after `gap` steps of the previous one (blocks join a
    running session), steps until everything completed.  Returns
    {ticket: (X, Y)} and the number of steps."""
    out = dict(out or {})
    order = list(out)
    queue = [(gap, list(items)) for gap, items in waves]
    steps = 0
    waiting = []
    active = 0
    while True:
        if queue and queue[0][0] <= 0:
            waiting.extend(queue.pop(0)[1])
        elif queue:
            queue[0] = (queue[0][0] - 1, queue[0][1])
        still = []
        for X, Y in waiting:
            t = sess.add(X, Y)
            if t is None:
                still.append((X, Y))
            else:
                out[t] = (X, Y)
                order.append(t)
        waiting = still
        inflight = sess.step()
        done, active, _ = sess.poll()
        for t in done:
            assert t in out
        if inflight >= 3:
            time.sleep(0.0002)
        steps += 1
        assert steps < max_steps, "session did not drain"
        if not queue and not waiting and active == 0:
            break
    torch.cuda.synchronize()
    return out, order, steps


@pytest.mark.parametrize("method", ["mixed", "mixed_mx", "composite_FP32"])
def test_session_joins_and_matches_alone(method):
    sizes = [64, 300, 1024, 33, 700, 128, 1, 512, 90, 2100, 256]
    xs, ps = _inputs(sizes)
    h = P.Handle(0, seed=5)
    sess = P.Session(h, 2 << 30, method=method)
    items = [(x, torch.empty_like(x)) for x in xs]
    # three waves: the later ones join while the first ones are mid-chain
    out, order, steps = _drive(sess, [(0, items[:4]), (5, items[4:8]), (7, items[8:])])
    print(f"session {method}: {len(out)} blocks in {steps} steps", flush=True)
    assert len(out) == len(sizes)
    ha = P.Handle(0, seed=5)
    cfg = P.default_config(method)
    for (x, y), pm, n in zip(items, ps, sizes):
        alone = P.project_batched(ha, [x], config=cfg)[0]
        torch.cuda.synchronize()
        assert torch.equal(alone, y), f"n={n} differs from its alone result"
        err = W.rel_fro_error(y, pm)
        assert err <= (1e-3 if method in ("mixed", "mixed_mx") else 2e-5), (n, err)
    sess.close()


def test_session_arena_limits_admission():
    # an arena with room for about two n=1024 blocks: later blocks are refused
    # (None) until completions free their regions, and all still finish
    sizes = [1024, 1024, 1024, 1024, 200, 1024]
    xs, ps = _inputs(sizes, seed0=950)
    h = P.Handle(0, seed=5)
    probe = P.Session(h, 64 << 20)
    per = probe.block_bytes(1024)
    probe.close()
    sess = P.Session(h, int(2.5 * per))
    items = [(x, torch.empty_like(x)) for x in xs]
    first = {sess.add(*it): it for it in items[:2]}
    assert None not in first
    assert sess.add(*items[2]) is None  # no room until one of the first two completes
    out, order, steps = _drive(sess, [(0, items[2:])], out=first)
    assert len(out) == len(sizes)
    ha = P.Handle(0, seed=5)
    for (x, y), pm in zip(items, ps):
        alone = P.project_batched(ha, [x])[0]
        torch.cuda.synchronize()
        assert torch.equal(alone, y)
    with pytest.raises(P.PsdpError) as e:
        sess.add(torch.zeros(4096, 4096, device=DEV), torch.zeros(4096, 4096, device=DEV))
    assert e.value.status == P.ERR_INVALID_ARG
    sess.close()
