"""GPU parity of the kernels behind the large-n projections.

The built-in tile choice runs 256 x 256 tiles from n = 4096, where the
oracle's emulation is too slow to compare with; here the 256 x 256 tiles are
forced at smaller n (psdp_set_tuned_tile), so that full projections run the
store, polynomial and (I + A)/2 epilogues of the 256 x 256 kernels (the
8-wave kernel of the two-piece formats, the 4-wave kernel of binary16), and
their result must be

  * bitwise the 64 x 64 tiling's result (same k order per entry), and
  * within the oracle's own error for the same method and input
    (0.7 e_oracle - 2e-6 <= e_gpu <= 1.3 e_oracle + 2e-6, as in
    test_gpu_parity.py), or, at sizes the emulation cannot reach, within the
    method's bar against the exact projection.

Then the benchmark workloads at full size (n = 8192, n = 16384, the 256-block
batch) against the closed form
Q max(L, 0) Q^T (generated on the device in float64), and the float32 Lanczos
bound that the projections use against the oracle's bound on the same
float32-rounded matrix and start vector (relative 1e-9).

Reference: src/composite_FP16.cu:106-233, src/composite_FP32.cu:89-156,
src/lanczos.cu:41-293, test/composite_FP32_test.hpp:80-129.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle import psdp_oracle as O
from _streams import start_vector

torch = pytest.importorskip("torch")
import psd_projection_amd as P  # noqa: E402
from psd_projection_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
GOLD = Path(__file__).resolve().parent / "golden"

# schedule of each method: (config overrides, formats whose products may take 256 x 256 tiles)
METHODS = {
    "mixed": (dict(method="mixed"), ["f16x3", "f16"]),
    "composite_FP32": (dict(method="composite_FP32"), ["f32"]),
    "bf16x3": (dict(method="custom", table="fp32", step_format=["bf16x3"] * 10, final_format="bf16x3"),
               ["bf16x3"]),
    "composite_FP32_emulated": (dict(method="composite_FP32_emulated"), ["bf16x9"]),
    "mixed_mx": (dict(method="mixed_mx"), ["f16mx", "f16"]),
}


def oracle_chain(Xs, name):
    c32, c16 = O.fp32_chain_coefficients(), O.fp16_chain_coefficients()
    if name == "mixed":
        return O.composite_chain(Xs, c16, ["f16x3"] * 4 + ["f16"] * 3, "f16", sym_mode="mirror")
    if name == "mixed_mx":
        return O.composite_chain(Xs, c16, ["f16mx"] * 4 + ["f16"] * 3, "f16", sym_mode="mirror")
    if name == "composite_FP32":
        return O.composite_chain(Xs, c32, ["f32"] * 10, "f32", sym_mode="mirror")
    if name == "bf16x3":
        return O.composite_chain(Xs, c32, ["bf16x3"] * 10, "bf16x3", sym_mode="mirror")
    return O.composite_chain(Xs, c32, ["bf16x9"] * 10, "bf16x9", sym_mode="mirror")


def project_tiled(handle, Xd, name, n, tile, seed):
    kw, fmts = METHODS[name]
    for f in fmts:
        handle.set_tile(n, f, tile)
    try:
        handle.set_seed(seed)
        return P.project(handle, Xd, **kw)
    finally:
        for f in fmts:
            handle.set_tile(n, f, 0)


@pytest.mark.parametrize("name", list(METHODS))
def test_big_tiles_projection_vs_oracle(handle, name):
    # n = 2304: 9 x 9 tiles of 256 (45 upper-triangle tiles, no tail launch)
    n, seed = 2304, 31
    print(f"big-tile projection {name} n={n}", flush=True)
    X, Q, lam = O.spectral_input(n, 3, "uniform_cluster")
    X32 = X.astype(np.float32)
    Xd = torch.from_numpy(X32).to(DEV)
    Y256 = project_tiled(handle, Xd, name, n, 256, seed)
    Y64 = project_tiled(handle, Xd, name, n, 64, seed)
    assert torch.equal(Y256, Y64)
    Y = Y256.cpu().numpy().astype(np.float64)
    assert np.array_equal(Y, Y.T)
    Pt = O.closed_form_projection(Q, lam)
    err = O.rel_fro_error(Y, Pt)
    _, up = O.lanczos_two_norm(X32.astype(np.float64), 20, 1e-10, q0=start_vector(seed, n))
    e_or = O.rel_fro_error(oracle_chain(X32.astype(np.float64) / up, name) * up, Pt)
    print(f"  gpu {err:.3e} oracle {e_or:.3e}", flush=True)
    assert 0.7 * e_or - 2e-6 <= err <= 1.3 * e_or + 2e-6
    if name in ("mixed", "mixed_mx"):
        assert err <= 1e-3


@pytest.mark.parametrize("n", [4100, 6000])
@pytest.mark.parametrize("name,bar", [("mixed", 1e-3), ("mixed_mx", 1e-3), ("bf16x3", 1e-4),
                                      ("composite_FP32", 1e-4)])
def test_big_tiles_projection_bitwise(handle, n, name, bar):
    # n = 4100: padded to 4352 = 17 x 256 (153 tiles, padding rows inside the
    # last tile row); n = 6000: 24 x 24 tiles = 300, more than one round on
    # 256 CUs, so the last partial round (44 tiles) runs either as 64 x 64
    # tiles in a second launch (split products off: bitwise the 64 x 64
    # tiling) or, by default, as 5 K chunks per tile plus their reduction,
    # whose epilogues (store, polynomial, (I + A)/2, output) must meet the
    # same bar.
    seed = 8
    print(f"big-tile projection {name} n={n}", flush=True)
    X, Pm, _ = W.spectral_matrix(n, 40 + n % 7, "uniform_cluster", device=DEV)
    Xd = X.to(torch.float32)
    del X
    handle.set_split_k(False)
    try:
        Y256 = project_tiled(handle, Xd, name, n, 256, seed)
    finally:
        handle.set_split_k(True)
    Y64 = project_tiled(handle, Xd, name, n, 64, seed)
    assert torch.equal(Y256, Y64)
    assert torch.equal(Y256, Y256.T)
    err = W.rel_fro_error(Y256, Pm)
    print(f"  gpu rel err {err:.3e}", flush=True)
    assert err <= bar
    Ys = project_tiled(handle, Xd, name, n, 256, seed)
    assert torch.equal(Ys, Ys.T)
    err_s = W.rel_fro_error(Ys, Pm)
    print(f"  split products: rel err {err_s:.3e}", flush=True)
    assert err_s <= bar
    if n == 4100:  # one round of 153 tiles: no split
        assert torch.equal(Ys, Y256)


@pytest.mark.parametrize("fmt", ["f16", "f16x3", "f16mx", "bf16x3", "e4m3"])
@pytest.mark.parametrize("n", [5700, 6144])
def test_split_products_match_whole_k(handle, n, fmt):
    # n = 5700: 23 x 23 tiles of 256 = 276, last round 20 tiles in 12 K
    # chunks; n = 6144: 300 tiles, last round 44 tiles in 5 chunks.  A split
    # tile sums its K in chunks and adds them in order: the same product to
    # float32 rounding, with every kernel of the 256 x 256 tiles (4-wave
    # one-piece, 8-wave two-piece, 8-wave F16MX) and the reduction's output
    # epilogue.
    g = torch.Generator(device=DEV).manual_seed(n + len(fmt))
    X = torch.randn(n, n, device=DEV, generator=g) / n ** 0.5
    X = (X + X.T) * 0.5
    Y = torch.randn(n, n, device=DEV, generator=g) / n ** 0.5
    Y = (Y + Y.T) * 0.5
    print(f"split product {fmt} n={n}", flush=True)
    handle.set_tile(n, fmt, 256)
    try:
        handle.set_split_k(False)
        C0 = P.sym_product(handle, X, Y, fmt)
        handle.set_split_k(True)
        C1 = P.sym_product(handle, X, Y, fmt)
    finally:
        handle.set_split_k(True)
        handle.set_tile(n, fmt, 0)
    torch.cuda.synchronize()
    d = float((C1.double() - C0.double()).norm() / C0.double().norm())
    print(f"  split vs whole K: {d:.2e}", flush=True)
    assert 0 < d <= 2e-6
    assert torch.equal(C1, C1.T)


# ------------------------------------------------------------------ full sizes
@pytest.mark.parametrize("method,bar", [("composite_FP32", 1e-4), ("composite_FP16", 5e-3), ("mixed", 1e-3),
                                        ("mixed_mx", 1e-3)])
def test_normal_cluster_n8192_full_size(handle, method, bar):
    # n = 8192, lambda ~ N(0,1) scaled to max 10 with 30% in [-1e-2, 1e-2].
    # The reference's FP16 chain is at 2-3e-3 on this spectrum (above 1e-3,
    # as its op order predicts); the FP32 chain and the shipped schedule stay
    # within their bars.
    n = 8192
    print(f"normal_cluster {method} n={n}", flush=True)
    X, Pm, _ = W.spectral_matrix(n, 2, "normal_cluster", device=DEV)
    Xd = X.to(torch.float32)
    del X
    handle.set_seed(2)
    Y = P.project(handle, Xd, method=method)
    err = W.rel_fro_error(Y, Pm)
    print(f"  rel err {err:.3e}", flush=True)
    assert err <= bar
    if method == "composite_FP16":
        assert err >= 1e-3  # the reference FP16 chain misses 1e-3 on this spectrum


N16384_SCHEDULES = {
    "mixed": (dict(method="mixed"), 1e-3),
    "mixed_mx": (dict(method="mixed_mx"), 1e-3),
    # the precision study's BF16x3 split products on the FP32 table (~6e-5)
    "bf16x3": (dict(method="custom", table="fp32", step_format=["bf16x3"] * 10, final_format="bf16x3"), 1e-4),
}


@pytest.mark.parametrize("name", list(N16384_SCHEDULES))
def test_uniform_cluster_n16384_full_size(handle, name):
    n = 16384
    kw, bar = N16384_SCHEDULES[name]
    print(f"uniform_cluster {name} n={n}", flush=True)
    X, Pm, _ = W.spectral_matrix(n, 3, "uniform_cluster", device=DEV)
    Xd = X.to(torch.float32)
    del X
    handle.set_seed(3)
    Y = P.project(handle, Xd, **kw)
    del Xd
    err = W.rel_fro_error(Y, Pm)
    print(f"  rel err {err:.3e}", flush=True)
    assert err <= bar


def test_e4m3_schedule_diverges_and_says_so(handle):
    # The precision study's FP8 schedule (e4m3 operands with a per-tensor
    # scale on the first half of the FP16-table steps, bfloat16 after) keeps
    # 3 mantissa bits: the first steps' coefficients (8.3, -23.1, 16.5)
    # amplify the ~6% operand error until iterates leave [-1, 1] and the
    # quintic diverges (the oracle's emulation of the schedule diverges too).
    # A synchronous call reports the non-finite result instead of returning it.
    n = 2048
    X, _, _ = W.spectral_matrix(n, 3, "uniform_cluster", device=DEV)
    Xd = X.to(torch.float32)
    del X
    cfg = P.default_config("custom", table="fp16", step_format=["e4m3"] * 3 + ["bf16"] * 4, final_format="bf16",
                           synchronize=1)
    handle.set_seed(3)
    with pytest.raises(P.PsdpError) as ei:
        P.project(handle, Xd, config=cfg)
    assert ei.value.status == P.ERR_NONFINITE


def test_256_block_batch_full(handle):
    # The benchmark's block-diagonal cone: 256 blocks, n log-uniform in
    # [32, 1024] (seed 4), lambda ~ U[-1, 1]; every block within 1e-3 of its
    # exact projection, for both shipped schedules.
    rng = np.random.default_rng(4)
    sizes = [int(round(float(np.exp(rng.uniform(np.log(32), np.log(1024)))))) for _ in range(256)]
    xs, ps = [], []
    for i, n in enumerate(sizes):
        X64, Pm, _ = W.spectral_matrix(n, 4000 + i, "uniform", device=DEV)
        xs.append(X64.float().contiguous())
        ps.append(Pm)
    torch.cuda.synchronize()
    for method in ("mixed_mx", "mixed"):
        print(f"batch of {len(sizes)} blocks, {method}", flush=True)
        outs = P.project_batched(handle, xs, method=method)
        errs = [W.rel_fro_error(y, pm) for y, pm in zip(outs, ps)]
        print(f"  max rel err {max(errs):.3e}", flush=True)
        assert max(errs) <= 1e-3
        for k in (0, 255):
            assert torch.equal(outs[k], outs[k].T)


# ------------------------------------------------------------------ Lanczos bound of the projections
@pytest.mark.parametrize("mode", ["reference", "tight"])
@pytest.mark.parametrize("n,kind", [(300, "normal"), (2100, "uniform_cluster"), (4100, "normal_cluster")])
def test_projection_bound_matches_oracle(handle, n, kind, mode):
    # The projections bound the float32-rounded matrix with the half-read
    # float32 GEMV (float64 accumulation); the oracle runs the same Lanczos in
    # float64 on the same float32 values and start vector.
    seed = 12
    X, _, lam = O.spectral_input(n, 5, kind)
    X32 = X.astype(np.float32)
    handle.set_seed(seed)
    P.project(handle, torch.from_numpy(X32).to(DEV), method="mixed", bound_mode=mode)
    lo, up, it = handle.last_bound()
    lo_o, up_o, st = O.lanczos_two_norm(X32.astype(np.float64), 20, 1e-10, q0=start_vector(seed, n),
                                        return_state=True, mode=mode)
    print(f"bound n={n} {kind} {mode}: up {up:.12g} oracle {up_o:.12g} iterations {it}", flush=True)
    assert it == st["nb"]
    assert lo == pytest.approx(lo_o, rel=1e-9)
    assert up == pytest.approx(up_o, rel=1e-9)
    nrm = np.abs(np.linalg.eigvalsh(X32.astype(np.float64))).max()
    assert lo <= nrm * (1 + 1e-9) <= up


@pytest.mark.parametrize("mode", ["reference", "tight"])
def test_batched_bounds_match_oracle(handle, mode):
    seed = 13
    sizes = [300, 77, 2100, 1]
    Xs = []
    for i, n in enumerate(sizes):
        X, _, _ = O.spectral_input(n, 60 + i, "normal")
        Xs.append(X.astype(np.float32))
    handle.set_seed(seed)
    P.project_batched(handle, [torch.from_numpy(x).to(DEV) for x in Xs], method="mixed", bound_mode=mode)
    got = handle.last_bounds_batched(len(sizes))
    for (lo, up, it), x, n in zip(got, Xs, sizes):
        lo_o, up_o = O.lanczos_two_norm(x.astype(np.float64), 20, 1e-10, q0=start_vector(seed, n), mode=mode)
        print(f"batched bound n={n} {mode}: up {up:.12g} oracle {up_o:.12g}", flush=True)
        assert lo == pytest.approx(lo_o, rel=1e-9)
        assert up == pytest.approx(up_o, rel=1e-9)
    with pytest.raises(P.PsdpError):
        handle.last_bounds_batched(len(sizes) + 1)


def test_tight_bound_projection_uniform_cluster(handle):
    # Opt-in tight bound: on the uniform_cluster spectrum (||X||_2 = 5) the
    # reference formula gives up ~ 7; the tight one stays within the Ritz
    # error of 5 and the projection keeps its accuracy.
    n = 2100
    X, Q, lam = O.spectral_input(n, 3, "uniform_cluster")
    Xd = torch.from_numpy(X.astype(np.float32)).to(DEV)
    Pt = O.closed_form_projection(Q, lam)
    handle.set_seed(4)
    Y_ref = P.project(handle, Xd, method="mixed")
    up_ref = handle.last_bound()[1]
    handle.set_seed(4)
    Y_t = P.project(handle, Xd, method="mixed", bound_mode="tight")
    up_t = handle.last_bound()[1]
    nrm = np.abs(lam).max()
    e_ref = O.rel_fro_error(Y_ref.cpu().numpy(), Pt)
    e_t = O.rel_fro_error(Y_t.cpu().numpy(), Pt)
    print(f"uniform_cluster n={n}: up {up_ref:.4f} (reference) {up_t:.4f} (tight), ||X|| {nrm:.4f}; "
          f"err {e_ref:.3e} / {e_t:.3e}", flush=True)
    assert nrm <= up_t < 1.01 * nrm < up_ref
    assert e_t <= 1e-3


# ------------------------------------------------------------------ transfer function fixtures
@pytest.mark.parametrize("table,fn,bar", [("fp32", P.composite_FP32, 1e-6), ("fp16", P.composite_FP16, 2e-3)])
def test_transfer_function_on_diagonal(handle, table, fn, bar):
    # The chain maps diag(x) to diag(g(x)), g(x) = x (1 + P(x)) / 2 with P the
    # composite polynomial of the table (tests/golden/transfer_<table>.npz,
    # float64): every product of diagonal matrices is elementwise, so the GPU
    # result is the scalar chain in float32 (binary16 operands for the FP16
    # method) and stays within its rounding of g.  Off-diagonal entries stay 0.
    d = np.load(GOLD / f"transfer_{table}.npz")
    x, g = d["x"], d["g"]
    n = x.size
    M = torch.diag(torch.from_numpy(x)).to(DEV, torch.float64)
    fn(handle, M)
    Y = M.cpu().numpy()
    assert np.count_nonzero(Y - np.diag(np.diag(Y))) == 0
    dev = np.abs(np.diag(Y) - g)
    print(f"transfer {table} n={n}: max |Y_ii - g(x_i)| = {dev.max():.3e}", flush=True)
    assert dev.max() <= bar
