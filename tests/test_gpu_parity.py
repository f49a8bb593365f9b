"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle.

Tolerances:
  * single products: relative Frobenius error vs the oracle's rounding-emulated
    product (same operand rounding, float32 accumulation) <= 2e-6;
  * projections: the reference's own test bars (rel <= 1e-3 / abs <= 0.1,
    test/composite_FP32_test.hpp) AND the oracle's own error for the same
    method and input within a stated margin (the GPU must reproduce the
    method's accuracy, not just pass 1e-3);
  * Lanczos (FP64 input, same start vector): lo/up to 1e-9 relative.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import psdp_oracle as O
from _streams import start_vector

torch = pytest.importorskip("torch")
import psd_projection_amd as P  # noqa: E402

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
KATS = json.loads((GOLD / "reference_kats.json").read_text())
DEV = "cuda:0"


def to_dev(M, dtype=torch.float64):
    # column-major n x n == transpose of the row-major array; matrices here
    # are symmetric, but keep the layout explicit for non-symmetric inputs
    return torch.from_numpy(np.asfortranarray(M).T.copy()).to(DEV, dtype)


def from_dev(T):
    return T.detach().cpu().numpy().astype(np.float64).T.copy()


# ------------------------------------------------------------------ products
@pytest.mark.parametrize("fmt,avg", [("f32", 0), ("f16", 0), ("f16x3", 0), ("bf16", 0), ("bf16x3", 0),
                                     ("bf16x9", 0), ("e4m3", 0), ("f16", 1), ("bf16", 1)])
@pytest.mark.parametrize("n", [64, 200, 1100, 2100])
def test_sym_product(handle, fmt, avg, n):
    # The kernel computes tiles with bi <= bj only; its result is the upper
    # triangle of X Y mirrored (or, averaged, (X Y + Y X)/2).  Compared with
    # the oracle's rounding-emulated product symmetrized the same way.
    print(f"sym_product fmt={fmt} avg={avg} n={n}", flush=True)
    X, _, _ = O.spectral_input(n, 100 + n, "uniform")
    X32 = X.astype(np.float32)
    Y32 = (X @ X).astype(np.float32)
    Y32 = ((Y32 + Y32.T) / 2).astype(np.float32)
    C = from_dev(P.sym_product(handle, to_dev(X32, torch.float32), to_dev(Y32, torch.float32), fmt,
                               average=bool(avg)))
    ref = O.sym_product(X32, Y32, fmt, "average" if avg else "mirror").astype(np.float64)
    assert np.array_equal(C, C.T)
    err = np.linalg.norm(C - ref) / np.linalg.norm(ref)
    assert err <= 2e-6, err


@pytest.mark.parametrize("fmt", ["f16", "f16x3", "bf16", "bf16x3"])
def test_sym_product_split_tail(handle, fmt):
    # n = 6144: 300 tiles of 256 x 256, i.e. one round on 256 CUs plus a partial
    # one of 44 tiles.  With split products off that round runs as 64 x 64
    # tiles in a second launch and the result is bitwise the 64 x 64 tiling's
    # (on devices with fewer CUs the rounds differ; the result must not); by
    # default its tiles run as 5 K chunks plus a reduction, within float32
    # rounding of the oracle's product.  The two-piece formats run their
    # 256 x 256 tiles on the 8-wave kernel, the one-piece formats on the
    # 4-wave kernel.
    n = 6144
    print(f"sym_product split tail fmt={fmt} n={n}", flush=True)
    X, _, _ = O.spectral_input(n, 77, "uniform")
    X32 = X.astype(np.float32)
    Xd = to_dev(X32, torch.float32)
    try:
        handle.set_tile(n, fmt, 256)
        Cs = from_dev(P.sym_product(handle, Xd, Xd, fmt))
        handle.set_split_k(False)
        C = from_dev(P.sym_product(handle, Xd, Xd, fmt))
        handle.set_tile(n, fmt, 64)
        C64 = from_dev(P.sym_product(handle, Xd, Xd, fmt))
    finally:
        handle.set_split_k(True)
        handle.set_tile(n, fmt, 0)
    assert np.array_equal(C, C64)
    ref = O.sym_product(X32, X32, fmt, "mirror").astype(np.float64)
    for M in (C, Cs):
        assert np.array_equal(M, M.T)
        err = np.linalg.norm(M - ref) / np.linalg.norm(ref)
        assert err <= 2e-6, err
    assert not np.array_equal(Cs, C)  # the split path ran


def test_sym_product_average_rejected_for_split_formats(handle):
    X = torch.eye(8, device=DEV)
    with pytest.raises(P.PsdpError):
        P.sym_product(handle, X, X, "f16x3", average=True)


# ------------------------------------------------------------------ KATs
def test_kat_deterministic_3x3(handle):
    k = KATS["composite_fp32_deterministic"]
    X = np.array(k["input_colmajor"]).reshape(3, 3, order="F")
    d = to_dev(X)
    P.composite_FP32(handle, d)
    E = np.array(k["expected_colmajor"]).reshape(3, 3, order="F")
    assert np.abs(from_dev(d) - E).max() <= k["tol"]


@pytest.mark.parametrize("fn,tol_scale", [(P.composite_FP32_auto_scale, 1.0),
                                          (P.composite_FP16_auto_scale, 5.0),
                                          (P.composite_FP32_emulated_auto_scale, 1.0)])
def test_kat_simple5(handle, fn, tol_scale):
    k = KATS["simple5"]
    d = to_dev(np.diag(k["diagonal"]))
    fn(handle, d)
    Y = from_dev(d)
    assert np.abs(Y - np.diag(k["expected_diagonal"])).max() <= k["tol"] * tol_scale


# ------------------------------------------------------------------ Lanczos
def test_lanczos_kat_5x5(handle):
    k = KATS["lanczos_5x5"]
    A = np.array(k["input_colmajor"]).reshape(5, 5, order="F")
    seed = 99
    handle.set_seed(seed)
    lo, up = P.approximate_two_norm(handle, to_dev(A), max_iter=k["max_iter"], tol=k["tol"])
    lo_o, up_o = O.lanczos_two_norm(A, k["max_iter"], k["tol"], q0=start_vector(seed, 5))
    assert lo <= k["two_norm"] + 1e-9 <= up
    assert lo == pytest.approx(lo_o, rel=1e-9)
    assert up == pytest.approx(up_o, rel=1e-9)


@pytest.mark.parametrize("n,kind,seed", [(300, "normal", 3), (1000, "normal_cluster", 4), (777, "uniform_cluster", 5)])
def test_lanczos_matches_oracle(handle, n, kind, seed):
    X, Q, lam = O.spectral_input(n, seed, kind)
    handle.set_seed(seed)
    lo, up = P.approximate_two_norm(handle, to_dev(X))
    lo_o, up_o = O.lanczos_two_norm(X, 20, 1e-10, q0=start_vector(seed, n))
    assert lo == pytest.approx(lo_o, rel=1e-9)
    assert up == pytest.approx(up_o, rel=1e-9)
    assert lo <= np.abs(lam).max() * (1 + 1e-12) <= up


def test_lanczos_zero_matrix(handle):
    lo, up = P.approximate_two_norm(handle, torch.zeros(50, 50, dtype=torch.float64, device=DEV))
    assert (lo, up) == (0.0, 1.0)


# ------------------------------------------------------------------ projections
def run_projection(handle, X, method, dtype=torch.float64, **kw):
    Xd = to_dev(X, dtype)
    Y = P.project(handle, Xd, method=method, **kw)
    return from_dev(Y)


def oracle_error(X, P_true, method, seed):
    """The oracle's error for the same method, with the library's
    symmetrization: upper triangle mirrored, except the FP16 method, which
    averages its non-square products like the reference's symmetrize."""
    n = X.shape[0]
    _, up = O.lanczos_two_norm(X, 20, 1e-10, q0=start_vector(seed, n))
    Xs = X / up
    c32, c16 = O.fp32_chain_coefficients(), O.fp16_chain_coefficients()
    if method == "composite_FP32":
        Y = O.composite_chain(Xs, c32, ["f32"] * 10, "f32", sym_mode="mirror")
    elif method == "composite_FP16":
        Y = O.composite_chain(Xs, c16, ["f16"] * 7, "f16", sym_mode="average")
    elif method == "composite_FP32_emulated":
        Y = O.composite_chain(Xs, c32, ["bf16x9"] * 10, "bf16x9", sym_mode="mirror")
    elif method == "mixed":
        Y = O.composite_chain(Xs, c16, ["f16x3"] * 4 + ["f16"] * 3, "f16", sym_mode="mirror")
    return O.rel_fro_error(Y * up, P_true)


def test_reference_uniform_scaled_1000(handle):
    # test/composite_FP32_test.hpp:80-129: n=1000, radius 1, no scaling
    A, Ppsd, _, _ = O.generate_and_project(1000, 1.0)
    d = to_dev(A)
    P.composite_FP32(handle, d)
    Y = from_dev(d)
    err = O.rel_fro_error(Y, Ppsd)
    e_or = O.rel_fro_error(O.composite_chain(A, O.fp32_chain_coefficients(), ["f32"] * 10, "f32",
                                             sym_mode="mirror"), Ppsd)
    print("uniform_scaled rel err gpu", err, "oracle", e_or)
    assert err <= 1e-3
    assert err <= 1.5 * e_or + 2e-6


@pytest.mark.parametrize("fn,method", [(P.composite_FP32_auto_scale, "composite_FP32"),
                                       (P.composite_FP32_emulated_auto_scale, "composite_FP32_emulated")])
def test_reference_uniform_autoscaled_1024(handle, fn, method):
    # test/composite_FP32_test.hpp:201-247: n=1024, radius 10, abs err <= 0.1
    A, Ppsd, _, _ = O.generate_and_project(1024, 10.0)
    seed = 17
    handle.set_seed(seed)
    d = to_dev(A)
    fn(handle, d)
    Y = from_dev(d)
    abs_err = np.linalg.norm(Y - Ppsd)
    err = abs_err / np.linalg.norm(Ppsd)
    e_or = oracle_error(A, Ppsd, method, seed)
    print(method, "autoscaled1024 abs", abs_err, "rel", err, "oracle rel", e_or)
    assert abs_err <= 0.1
    assert err <= 1.5 * e_or + 2e-6


@pytest.mark.parametrize("method,kind,n,bar", [
    ("composite_FP32", "uniform", 512, 2e-5),
    ("composite_FP32_emulated", "uniform", 512, 2e-5),
    ("composite_FP16", "normal_cluster", 512, None),
    ("composite_FP16", "uniform", 512, None),
    ("mixed", "normal_cluster", 1024, 1e-3),
    ("mixed", "uniform_cluster", 1024, 1e-3),
    ("mixed", "uniform", 640, 1e-3),
    ("mixed", "uniform_cluster", 2100, 1e-3),
])
def test_projection_accuracy(handle, method, kind, n, bar):
    print(f"projection {method} {kind} n={n}", flush=True)
    X, Q, lam = O.spectral_input(n, 1 if kind == "uniform" else (2 if kind == "normal_cluster" else 3), kind)
    Ptrue = O.closed_form_projection(Q, lam)
    seed = 5
    handle.set_seed(seed)
    Y = run_projection(handle, X, method)
    err = O.rel_fro_error(Y, Ptrue)
    e_or = oracle_error(X, Ptrue, method, seed)
    print(f"  gpu rel err {err:.3e} oracle {e_or:.3e}")
    assert np.array_equal(Y, Y.T)
    if bar is not None:
        assert err <= bar
    # same method, same rounding: within 30% (+ float32 accumulation noise)
    assert err <= 1.3 * e_or + 2e-6
    assert err >= 0.7 * e_or - 2e-6


def test_n512_fp32_io(handle):
    # n=512, FP32 storage, composite_FP32 with the Lanczos bound
    X, Q, lam = O.uniform512_input(512)
    Ptrue = O.closed_form_projection(Q, lam)
    Y = run_projection(handle, X, "composite_FP32", dtype=torch.float32)
    err = O.rel_fro_error(Y, Ptrue)
    print("n512 fp32 rel err", err)
    assert err <= 2e-5


# ------------------------------------------------------------------ edges
def test_zero_matrix(handle):
    Z = torch.zeros(70, 70, dtype=torch.float64, device=DEV)
    for m in ("composite_FP32", "composite_FP16", "mixed", "composite_FP32_emulated"):
        Y = P.project(handle, Z, method=m)
        assert torch.count_nonzero(Y).item() == 0


@pytest.mark.parametrize("val", [-3.0, 2.0, 0.25])
def test_one_by_one(handle, val):
    d = torch.tensor([[val]], dtype=torch.float64, device=DEV)
    P.composite_FP32_auto_scale(handle, d)
    assert d.item() == pytest.approx(max(val, 0.0), abs=1e-4 * abs(val))


def test_two_by_two(handle):
    X = np.array([[1.0, 2.0], [2.0, -1.0]])
    w, V = np.linalg.eigh(X)
    expect = (V * np.maximum(w, 0)) @ V.T
    Y = run_projection(handle, X, "composite_FP32")
    assert np.abs(Y - expect).max() < 1e-4


def three_by_three_undercut():
    # A 3 x 3 input whose bound under the reference's convergence rule
    # (src/lanczos.cu:145 drops the third, Krylov-space-completing iteration)
    # undercuts ||X||_2 with the start vector of seed 9 (see test_oracle.py)
    rng = np.random.default_rng(148)
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    lam = rng.uniform(-5.0, 5.0, 3)
    X = (Q * lam) @ Q.T
    return 0.5 * (X + X.T), Q, lam


@pytest.mark.parametrize("method", ["mixed_mx", "composite_FP32"])
def test_three_by_three_bound_covers_norm(handle, method):
    X, Q, lam = three_by_three_undercut()
    handle.set_seed(9)
    print(f"3x3 bound {method}", flush=True)
    Xd = torch.tensor(X, dtype=torch.float32, device=DEV)
    Y = P.project(handle, Xd, method=method)
    lo, up, it = handle.last_bound()
    Xf = X.astype(np.float32).astype(np.float64)
    lo_o, up_o = O.lanczos_two_norm(Xf, 20, 1e-10, q0=start_vector(9, 3))
    nrm = np.linalg.norm(Xf, 2)
    assert it == 3 and up == pytest.approx(up_o, rel=1e-9) and up >= nrm
    assert lo == pytest.approx(nrm, rel=1e-9)
    Pm = (Q * np.maximum(lam, 0.0)) @ Q.T
    err = np.linalg.norm(Y.double().cpu().numpy() - Pm) / np.linalg.norm(Pm)
    assert err <= (1e-3 if method == "mixed_mx" else 1e-4), err


@pytest.mark.parametrize("method,bar", [("mixed_mx", 1.5e-3), ("composite_FP32", 1e-4)])
def test_tiny_sizes(handle, method, bar):
    # n = 1 .. 21: the bound's Krylov space completes (n <= max_iter), the
    # tiles are mostly padding; error against ||X||_F (the positive part of a
    # tiny spectrum can vanish), U[-5, 5] spectra with Haar bases
    handle.set_seed(9)
    print(f"tiny sizes {method}", flush=True)
    for n in range(1, 22):
        X, Q, lam = O.spectral_input(n, 100 + n, "uniform_cluster")
        Y = P.project(handle, torch.tensor(X, dtype=torch.float32, device=DEV), method=method)
        Pm = (Q * np.maximum(lam, 0.0)) @ Q.T
        e = np.linalg.norm(Y.double().cpu().numpy() - Pm) / np.linalg.norm(X)
        assert e <= bar, (n, e)


def test_nonfinite_input_reported(handle):
    X = torch.eye(40, dtype=torch.float64, device=DEV)
    X[3, 5] = float("nan")
    with pytest.raises(P.PsdpError) as e:
        P.composite_FP32_auto_scale(handle, X)
    assert e.value.status == P.ERR_NONFINITE


def test_in_place_equals_out_of_place_and_deterministic(handle):
    X, _, _ = O.spectral_input(300, 9, "normal")
    handle.set_seed(3)
    a = P.project(handle, to_dev(X), method="mixed")
    b = to_dev(X)
    handle.set_seed(3)
    P.project(handle, b, out=b, method="mixed")
    handle.set_seed(3)
    c = P.project(handle, to_dev(X), method="mixed")
    assert torch.equal(a, b) and torch.equal(a, c)


def test_explicit_workspace(handle):
    X, Q, lam = O.spectral_input(256, 4, "uniform")
    need = P.workspace_size(256, "composite_FP32")
    ws = torch.empty(need, dtype=torch.uint8, device=DEV)
    handle.set_seed(1)
    Y1 = P.project(handle, to_dev(X), method="composite_FP32", workspace=ws)
    handle.set_seed(1)
    Y2 = P.project(handle, to_dev(X), method="composite_FP32")
    assert torch.equal(Y1, Y2)
    with pytest.raises(P.PsdpError):
        P.project(handle, to_dev(X), method="composite_FP32", workspace=ws[: need // 2])


# ------------------------------------------------------------------ batched
def batched_inputs(count, seed, lo=32, hi=300):
    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(int)
    out = []
    for i, n in enumerate(ns):
        X, Q, lam = O.spectral_input(int(n), seed * 1000 + i, "uniform")
        out.append((X, O.closed_form_projection(Q, lam)))
    return out


@pytest.mark.parametrize("method,bar", [("mixed", 1e-3), ("composite_FP32", 2e-5)])
def test_batched_accuracy_and_batch_invariance(handle, method, bar):
    data = batched_inputs(24, 4)
    Xs = [to_dev(X, torch.float32) for X, _ in data]
    handle.set_seed(11)
    Ys = P.project_batched(handle, Xs, method=method)
    worst = 0.0
    for (X, Pt), Y in zip(data, Ys):
        err = O.rel_fro_error(from_dev(Y), Pt)
        worst = max(worst, err)
        assert err <= bar, (X.shape[0], err)
    print(f"batched {method}: worst block rel err {worst:.3e}")
    # the same block alone (single-matrix path) gives bitwise the same result
    for k in (0, 7, 23):
        handle.set_seed(11)
        alone = P.project(handle, Xs[k], method=method)
        assert torch.equal(alone, Ys[k]), k
    # and in a different mix (other blocks, other order)
    handle.set_seed(11)
    mix = [Xs[23]] + [to_dev(X, torch.float32) for X, _ in batched_inputs(5, 9)] + [Xs[0]]
    Ym = P.project_batched(handle, mix, method=method)
    assert torch.equal(Ym[0], Ys[23]) and torch.equal(Ym[-1], Ys[0])


@pytest.mark.parametrize("method,bar", [("mixed", 1.5e-3), ("composite_FP32", 1e-4)])
def test_batched_tiny_blocks(handle, method, bar):
    # the small cones of an SDP (n = 1 .. 40, U[-5, 5] spectra): every block
    # within the method's error of ||X||_F, and bitwise its result alone
    sizes = list(range(1, 41))
    data = [O.spectral_input(n, 600 + n, "uniform_cluster") for n in sizes]
    Xs = [torch.tensor(X, dtype=torch.float32, device=DEV) for X, _, _ in data]
    handle.set_seed(11)
    print(f"batched tiny blocks {method}", flush=True)
    Ys = P.project_batched(handle, Xs, method=method)
    for (X, Q, lam), Y in zip(data, Ys):
        Pm = (Q * np.maximum(lam, 0.0)) @ Q.T
        e = np.linalg.norm(Y.double().cpu().numpy() - Pm) / np.linalg.norm(X)
        assert e <= bar, (X.shape[0], e)
    for k in (0, 2, 19, 39):
        handle.set_seed(11)
        assert torch.equal(P.project(handle, Xs[k], method=method), Ys[k]), sizes[k]


def test_batched_rejects_bad_blocks(handle):
    with pytest.raises(P.PsdpError):
        P.project_batched(handle, [])


def test_packed_cone_equals_block_list(handle):
    """The packed block-diagonal layout (blocks back to back in one buffer,
    projected in place or out of place) gives bitwise the block-list result."""
    data = batched_inputs(12, 6, lo=20, hi=400)
    Xs = [to_dev(X, torch.float32) for X, _ in data]
    sizes = [x.shape[0] for x in Xs]
    handle.set_seed(5)
    ref = P.project_batched(handle, Xs)
    packed = torch.cat([x.reshape(-1) for x in Xs])
    keep = packed.clone()
    handle.set_seed(5)
    out = P.project_packed(handle, packed, sizes, out=torch.empty_like(packed))
    assert torch.equal(packed, keep)  # out of place leaves the input alone
    handle.set_seed(5)
    inplace = P.project_packed(handle, packed, sizes)
    assert inplace.data_ptr() == packed.data_ptr()
    off = 0
    for n, y in zip(sizes, ref):
        assert torch.equal(out[off:off + n * n].view(n, n), y)
        assert torch.equal(packed[off:off + n * n].view(n, n), y)
        off += n * n
    with pytest.raises(ValueError):
        P.project_packed(handle, packed, sizes[:-1])
    with pytest.raises(P.PsdpError):
        P.project_packed(handle, packed[:4], [0, 2])


# ------------------------------------------------------------------ EVD drop-ins
def test_eig_fp64_psd(handle):
    X, Q, lam = O.spectral_input(300, 21, "normal")
    d = to_dev(X)
    w = P.eig_FP64_psd(handle, d, return_eigenvalues=True)
    assert O.rel_fro_error(from_dev(d), O.closed_form_projection(Q, lam)) < 1e-12
    assert np.allclose(w.cpu().numpy(), np.sort(lam), atol=1e-12)
    k = KATS["simple5"]
    d5 = to_dev(np.diag(k["diagonal"]))
    assert P.eig_FP64_psd(handle, d5) is None
    assert np.abs(from_dev(d5) - np.diag(k["expected_diagonal"])).max() < 1e-14


def test_eig_fp32_psd(handle):
    X, Q, lam = O.spectral_input(200, 22, "normal")
    d = to_dev(X)
    w = P.eig_FP32_psd(handle, d, return_eigenvalues=True)
    assert O.rel_fro_error(from_dev(d), O.closed_form_projection(Q, lam)) < 2e-5
    assert np.allclose(w.cpu().numpy(), np.sort(lam), atol=1e-4)


# ------------------------------------------------------------------ diverged chain
@pytest.mark.parametrize("fn", [P.composite_FP32, P.composite_FP16, P.composite_FP32_emulated])
def test_unscaled_divergence_reported(handle, fn):
    # include/psd_projection/composite_FP32.h:11: without the bound a spectrum
    # outside [-1, 1] makes the chain diverge (p_1(3) ~ 3e3, then overflow);
    # the output epilogue sees the non-finite values and the call says so.
    d = to_dev(np.diag([3.0, -2.0]))
    with pytest.raises(P.PsdpError) as e:
        fn(handle, d)
    assert e.value.status == P.ERR_NONFINITE
    # the same matrix through the auto-scaled entry point is fine
    d = to_dev(np.diag([3.0, -2.0]))
    P.composite_FP32_auto_scale(handle, d)
    assert np.abs(from_dev(d) - np.diag([3.0, 0.0])).max() < 1e-4


def test_unscaled_nan_input_reported(handle):
    X = torch.eye(40, dtype=torch.float64, device=DEV) * 0.5
    X[3, 5] = X[5, 3] = float("nan")
    with pytest.raises(P.PsdpError) as e:
        P.composite_FP32(handle, X)
    assert e.value.status == P.ERR_NONFINITE
    # a later good call on the same handle is not affected by the old flag
    Y = torch.eye(40, dtype=torch.float64, device=DEV) * 0.5
    P.composite_FP32(handle, Y)
    assert torch.allclose(Y, torch.eye(40, dtype=torch.float64, device=DEV) * 0.5, atol=1e-6)


# ------------------------------------------------------------------ EVD drop-ins, reference ownership
def _read_device_doubles(ptr, n):
    """Copies n doubles at a raw device address to the host (hipMemcpy)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    out = np.empty(n, dtype=np.float64)
    st = hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr), ctypes.c_size_t(8 * n), 2)
    assert st == 0, st  # hipMemcpyDeviceToHost
    return out


@pytest.mark.parametrize("fp64", [True, False])
def test_eig_psd_returns_owned_eigenvalues(handle, fp64):
    # src/eig_FP64_psd.cu:46-49,84: with return_eigenvalues the call returns a
    # device array the caller frees (MATLAB/psd_projection_MATLAB.cu:158-162,
    # 175-179); without it nothing is allocated (null, never uninitialised).
    X, Q, lam = O.spectral_input(150, 23, "normal")
    d = to_dev(X)
    ptr = P.eig_psd_alloc(handle, d, fp64=fp64, return_eigenvalues=True)
    assert ptr != 0
    w = _read_device_doubles(ptr, 150)
    P.free_device(ptr)
    tol = 1e-12 if fp64 else 1e-4
    assert np.allclose(w, np.sort(lam), atol=tol)
    assert O.rel_fro_error(from_dev(d), O.closed_form_projection(Q, lam)) < (1e-12 if fp64 else 2e-5)
    d2 = to_dev(X)
    assert P.eig_psd_alloc(handle, d2, fp64=fp64, return_eigenvalues=False) == 0
    assert torch.equal(d, d2)


# ------------------------------------------------------------------ method strings (MEX dispatch)
@pytest.mark.parametrize("method,bar", [("composite_FP32", 2e-5), ("composite_FP16", 5e-3),
                                        ("composite_FP32_emulated", 2e-5), ("eig_FP64", 1e-12),
                                        ("eig_FP32", 2e-5), ("mixed", 1e-3)])
def test_method_strings(handle, method, bar):
    # MATLAB/psd_projection_MATLAB.cu:112-166: one entry point keyed by the
    # method string, in place on a float64 matrix; the eig_* methods also
    # return the eigenvalues.
    X, Q, lam = O.spectral_input(400, 24, "uniform")
    Pt = O.closed_form_projection(Q, lam)
    d = to_dev(X)
    _, w = P.project_method(handle, d, method, return_eigenvalues=True)
    err = O.rel_fro_error(from_dev(d), Pt)
    print(f"method {method}: rel err {err:.3e}", flush=True)
    assert err <= bar
    if method.startswith("eig_"):
        assert np.allclose(w.cpu().numpy(), np.sort(lam), atol=1e-4)
    else:
        assert w is None
    # the float32 front end takes the same names
    Y = P.project(handle, to_dev(X, torch.float32), method=method)
    assert O.rel_fro_error(from_dev(Y), Pt) <= max(bar, 2e-5)
    with pytest.raises(P.PsdpError) as e:
        P.project_method(handle, d, "eig_FP16")
    assert e.value.status == P.ERR_INVALID_ARG


# ------------------------------------------------------------------ stream changes
def test_two_streams_one_handle_ordered(handle):
    # Two asynchronous projections through one handle on two torch streams:
    # they share the handle's workspace, so the second stream must wait for the
    # first one's work (psdp_set_stream orders them); both results equal the
    # synchronous ones bitwise.
    X1, _, _ = O.spectral_input(900, 41, "normal")
    X2, _, _ = O.spectral_input(700, 42, "uniform")
    a, b = to_dev(X1, torch.float32), to_dev(X2, torch.float32)
    handle.set_seed(9)
    ref1 = P.project(handle, a, method="mixed")
    handle.set_seed(9)
    ref2 = P.project(handle, b, method="mixed")
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(3):
        o1, o2 = torch.empty_like(a), torch.empty_like(b)
        handle.set_seed(9)
        with torch.cuda.stream(s1):
            P.project(handle, a, out=o1, method="mixed", synchronize=0)
        with torch.cuda.stream(s2):
            P.project(handle, b, out=o2, method="mixed", synchronize=0)
        torch.cuda.synchronize()
        assert torch.equal(o1, ref1) and torch.equal(o2, ref2)
