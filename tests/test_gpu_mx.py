"""FMT_F16MX on the GPU: binary16 hi.hi products plus the cross terms
hi.lo + lo.hi from MX-scaled e4m3 copies on the scaled FP8 MFMA.

* Products through psdp_sym_product_f32 (pieces written by the scale/split
  kernel) against the oracle's emulation (oracle.mx_e4m3 blocks, float32
  sums) for every tile kernel: 64 x 64 and 128 x 128 four-wave tiles, the
  8-wave 256 x 256 tile and its 64 x 64 tail launch.  The pieces are the
  same values on both sides (the e4m3 rounding of v_cvt_pk_fp8_f32 matches
  round_e4m3), so the results differ by float32 summation order only.
* The mixed_mx schedule (its epilogues write F16MX pieces) in batched calls
  and sessions: bitwise its result alone, within 1e-3 of the exact
  projection.  Full projections on the 256 x 256 kernels against the oracle
  are in test_gpu_headline.py.

No reference counterpart (the library's own operand format); the oracle
functions used here are pinned in test_oracle.py.
"""
import numpy as np
import pytest

from oracle import psdp_oracle as O

torch = pytest.importorskip("torch")
import psd_projection_amd as P  # noqa: E402
from psd_projection_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _operand(n, seed, kind):
    X, _, lam = O.spectral_input(n, seed, kind)
    X = X / np.abs(lam).max()
    return X.astype(np.float32)


@pytest.mark.parametrize("n,tile", [(200, 64), (1000, 64), (1100, 128), (1100, 64), (2304, 256), (4100, 256)])
def test_f16mx_product_matches_oracle(handle, n, tile):
    # n = 4100 pads to 17 x 256: padding inside the last tile row, and the
    # last partial round of 256 x 256 tiles runs as 64 x 64 tiles
    print(f"f16mx product n={n} tile={tile}", flush=True)
    X, Y = _operand(n, 70, "uniform"), _operand(n, 71, "normal")
    handle.set_tile(n, "f16mx", tile)
    try:
        C = P.sym_product(handle, torch.from_numpy(X).to(DEV), torch.from_numpy(Y).to(DEV), "f16mx")
    finally:
        handle.set_tile(n, "f16mx", 0)
    C = C.cpu().numpy().astype(np.float64)
    assert np.array_equal(C, C.T)
    # the kernels compute the upper triangle of X Y and mirror it (X and Y do
    # not commute here, so X Y itself is not symmetric)
    mirror = lambda m: np.triu(m) + np.triu(m, 1).T
    ex = mirror(X.astype(np.float64) @ Y.astype(np.float64))
    ref = mirror(O.emulated_product(X, Y, "f16mx").astype(np.float64))
    nrm = np.linalg.norm(ex)
    e_gpu, e_or = np.linalg.norm(C - ex) / nrm, np.linalg.norm(ref - ex) / nrm
    d = np.linalg.norm(C - ref) / nrm
    print(f"  gpu {e_gpu:.3e} oracle {e_or:.3e} |gpu - oracle| {d:.3e}", flush=True)
    assert d <= 1e-6
    assert 0.8 * e_or <= e_gpu <= 1.25 * e_or
    # between FP16x3 (~3e-7) and binary16 (~3e-4)
    assert 2e-6 <= e_gpu <= 5e-5


def test_f16mx_product_is_tile_independent(handle):
    n = 2304
    X, Y = _operand(n, 72, "normal"), _operand(n, 73, "uniform")
    Xd, Yd = torch.from_numpy(X).to(DEV), torch.from_numpy(Y).to(DEV)
    out = {}
    for tile in (64, 128, 256):
        handle.set_tile(n, "f16mx", tile)
        try:
            out[tile] = P.sym_product(handle, Xd, Yd, "f16mx")
        finally:
            handle.set_tile(n, "f16mx", 0)
    torch.cuda.synchronize()
    assert torch.equal(out[64], out[128]) and torch.equal(out[64], out[256])


def test_mixed_mx_batched_equals_alone(handle):
    sizes = [64, 300, 1024, 33, 700, 1, 2100]
    xs, pms = [], []
    for i, n in enumerate(sizes):
        X64, Pm, _ = W.spectral_matrix(n, 500 + i, "normal")
        xs.append(X64.float().contiguous())
        pms.append(Pm)
    torch.cuda.synchronize()
    print(f"mixed_mx batched sizes={sizes}", flush=True)
    handle.set_seed(21)
    outs = P.project_batched(handle, xs, method="mixed_mx")
    torch.cuda.synchronize()
    for x, y, pm, n in zip(xs, outs, pms, sizes):
        handle.set_seed(21)
        alone = P.project_batched(handle, [x], method="mixed_mx")[0]
        torch.cuda.synchronize()
        assert torch.equal(alone, y), n
        err = W.rel_fro_error(y, pm)
        assert err <= 1e-3, (n, err)


def test_mixed_mx_session_equals_batched(handle):
    sizes = [256, 64, 1024, 500]
    xs = []
    for i, n in enumerate(sizes):
        X64, _, _ = W.spectral_matrix(n, 540 + i, "uniform")
        xs.append(X64.float().contiguous())
    torch.cuda.synchronize()
    h = P.Handle(0, seed=5)
    sess = P.Session(h, 1 << 30, method="mixed_mx")
    items = [(x, torch.empty_like(x)) for x in xs]
    pending = {}
    for x, y in items[:2]:
        pending[sess.add(x, y)] = (x, y)
    steps, added = 0, 2
    while True:
        sess.step()
        steps += 1
        if steps == 5:  # the rest join mid-chain
            for x, y in items[2:]:
                pending[sess.add(x, y)] = (x, y)
            added = len(items)
        done, active, _ = sess.poll()
        if added == len(items) and active == 0 and steps > 5:
            break
        assert steps < 5000
    torch.cuda.synchronize()
    ha = P.Handle(0, seed=5)
    for x, y in items:
        alone = P.project_batched(ha, [x], method="mixed_mx")[0]
        torch.cuda.synchronize()
        assert torch.equal(alone, y)
    sess.close()


# Alternating binary16x3 and F16MX steps: the binary16x3 polynomial epilogue
# writing F16MX pieces (A^3 read back from its binary16x3 pieces, A from
# float32) and the F16MX epilogue writing binary16x3 pieces, on the 64 x 64
# kernels and (n = 2304) on the 8-wave 256 x 256 kernels, bitwise the same.
ALT_STEPS = ["f16x3", "f16mx", "f16x3", "f16mx", "f16", "f16", "f16"]
ALT = dict(method="custom", table="fp16", step_format=ALT_STEPS, final_format="f16")


@pytest.mark.parametrize("n", [600, 2304])
def test_alternating_f16x3_f16mx_schedule(handle, n):
    from _streams import start_vector
    seed = 41
    print(f"alternating f16x3/f16mx schedule n={n}", flush=True)
    X, Q, lam = O.spectral_input(n, 6, "uniform_cluster")
    X32 = X.astype(np.float32)
    Xd = torch.from_numpy(X32).to(DEV)
    fmts = ("f16x3", "f16mx", "f16")

    def run(tile):
        for f in fmts:
            handle.set_tile(n, f, tile)
        try:
            handle.set_seed(seed)
            return P.project(handle, Xd, **ALT)
        finally:
            for f in fmts:
                handle.set_tile(n, f, 0)

    Y = run(64)
    if n >= 2304:
        # 9 x 9 tiles of 256, one partial round: every tiling sums each entry
        # in the same order
        assert torch.equal(run(256), Y)
    Yn = Y.cpu().numpy().astype(np.float64)
    assert np.array_equal(Yn, Yn.T)
    Pt = O.closed_form_projection(Q, lam)
    err = O.rel_fro_error(Yn, Pt)
    _, up = O.lanczos_two_norm(X32.astype(np.float64), 20, 1e-10, q0=start_vector(seed, n))
    c16 = O.fp16_chain_coefficients()
    Yo = O.composite_chain(X32.astype(np.float64) / up, c16, ALT_STEPS, "f16", sym_mode="mirror") * up
    e_or = O.rel_fro_error(Yo, Pt)
    print(f"  gpu {err:.3e} oracle {e_or:.3e}", flush=True)
    assert 0.7 * e_or - 2e-6 <= err <= 1.3 * e_or + 2e-6
    assert err <= 1e-3


def test_all_f16x3_schedule_half_identity_epilogue(handle):
    # every step binary16x3 and a binary16x3 final product: the last
    # polynomial epilogue ((I + A)/2) also reads A^3 back from its pieces
    from _streams import start_vector
    n, seed = 500, 43
    steps = ["f16x3"] * 7
    X, Q, lam = O.spectral_input(n, 7, "uniform_cluster")
    X32 = X.astype(np.float32)
    handle.set_seed(seed)
    Y = P.project(handle, torch.from_numpy(X32).to(DEV), method="custom", table="fp16", step_format=steps,
                  final_format="f16x3")
    Yn = Y.cpu().numpy().astype(np.float64)
    Pt = O.closed_form_projection(Q, lam)
    err = O.rel_fro_error(Yn, Pt)
    _, up = O.lanczos_two_norm(X32.astype(np.float64), 20, 1e-10, q0=start_vector(seed, n))
    Yo = O.composite_chain(X32.astype(np.float64) / up, O.fp16_chain_coefficients(), steps, "f16x3",
                           sym_mode="mirror") * up
    e_or = O.rel_fro_error(Yo, Pt)
    print(f"all-f16x3 schedule n={n}: gpu {err:.3e} oracle {e_or:.3e}", flush=True)
    assert 0.7 * e_or - 2e-6 <= err <= 1.3 * e_or + 2e-6
    assert err <= 1e-3
