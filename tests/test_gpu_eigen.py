"""GPU parity of the eigenpair routines and the deflated projections (through
the C ABI) against the oracle and the reference's own test bars.

Tolerances:
  * LOBPCG eigenvalues: the reference's bars (1e-4 on the 6x6 KAT,
    lobpcg_test.hpp:46; 1e-8 against the FP64 EVD, lobpcg_test.hpp:110);
  * compute_eigenpairs on short runs (orthogonal basis, same start vector as
    the oracle): values and bound to 1e-10 relative, vectors to 1e-8 up to
    sign; on full-length runs (basis orthogonality lost, where rounding
    differences grow): each kept pair is an eigenpair of A within its
    residual filter, kept vectors overlap <= ortho_tol, bound >= ||A||;
  * deflated projections: the reference's bars (composite_FP32_test.hpp:283,
    :329; composite_FP16_test.hpp:57, :129) and the oracle's error for the
    same input within a stated margin.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import psdp_oracle as O
from _streams import start_block, start_vector

torch = pytest.importorskip("torch")
import psd_projection_amd as P  # noqa: E402

pytestmark = pytest.mark.gpu
KATS = json.loads((Path(__file__).resolve().parent / "golden" / "reference_kats.json").read_text())
DEV = "cuda:0"


def to_dev(M):
    return torch.from_numpy(np.asfortranarray(M).T.copy()).to(DEV, torch.float64)


def from_dev(T):
    return T.detach().cpu().numpy().astype(np.float64).T.copy()


def rand_symmetric(n, seed=1):
    g = O.GlibcRandom(seed)
    M = g.take(n * n).astype(np.float64).reshape(n, n, order="F") / O.GlibcRandom.RAND_MAX
    return M + M.T


# ------------------------------------------------------------------ LOBPCG
def test_lobpcg_kat_6x6(handle):
    k = KATS["lobpcg_6x6"]
    A = np.array(k["input_colmajor"]).reshape(6, 6, order="F")
    print("lobpcg 6x6 KAT", flush=True)
    V, D = P.lobpcg(handle, to_dev(A), 2, maxiter=100, tol=1e-8)
    d = D.cpu().numpy()
    assert np.allclose(d, k["expected_top2_reference"], atol=k["tol"])
    assert np.allclose(d, sorted(k["eigenvalues"])[::-1][:2], atol=1e-10)
    Vh = V.cpu().numpy()
    assert np.allclose(Vh.T @ Vh, np.eye(2), atol=1e-12)
    assert np.linalg.norm(A @ Vh - Vh * d) <= 1e-8


@pytest.mark.parametrize("n,m", [(100, 10), (1000, 100)])
def test_lobpcg_random_matches_evd_and_oracle(handle, n, m):
    # lobpcg_test.hpp:57-121 (rand()/RAND_MAX entries, A + A^T, m = n/10)
    A = rand_symmetric(n)
    seed = 77
    handle.set_seed(seed)
    print(f"lobpcg random n={n} m={m}", flush=True)
    V, D = P.lobpcg(handle, to_dev(A), m)
    it, rn = handle.lobpcg_last()
    d = D.cpu().numpy()
    ev = np.linalg.eigvalsh(A)[::-1][:m]
    assert np.abs(d - ev).max() <= 1e-8
    _, lam_o, it_o, rn_o = O.lobpcg(A, m, start_block(seed, n, m), 100, 1e-8)
    assert np.abs(d - lam_o).max() <= 1e-9
    # iterations to reach tol depend on rounding once the dX block is tiny
    assert abs(it - it_o) <= max(5, it_o // 5), (it, it_o)
    Vh = V.cpu().numpy()
    assert np.abs(Vh.T @ Vh - np.eye(m)).max() <= 1e-10


def test_lobpcg_warm_start(handle):
    A = rand_symmetric(300, seed=5)
    print("lobpcg warm start n=300 m=20", flush=True)
    V, D = P.lobpcg(handle, to_dev(A), 20, maxiter=300, tol=1e-9)
    V2, D2 = P.lobpcg(handle, to_dev(A), 20, V=V.clone(), warmstart=True, maxiter=300, tol=1e-8)
    it, rn = handle.lobpcg_last()
    assert it == 0 and rn < 1e-8
    assert torch.allclose(D2, D, atol=1e-12)


def test_lobpcg_rejects_3m_above_n(handle):
    A = to_dev(np.eye(10))
    with pytest.raises(P.PsdpError):
        P.lobpcg(handle, A, 4)


# ------------------------------------------------------------------ Lanczos eigenpairs
@pytest.mark.parametrize("row", ["first", "last"])
def test_eigenpairs_kat_5x5(handle, row):
    k = KATS["eigenpairs_5x5"]
    A = np.array(k["input_colmajor"]).reshape(5, 5, order="F")
    seed = 31
    handle.set_seed(seed)
    handle.set_eigenpair_options(row)
    print(f"eigenpairs 5x5 row={row}", flush=True)
    vals, vecs, bound = P.compute_eigenpairs(handle, to_dev(A), k["k"])
    handle.set_eigenpair_options("first")
    assert bound >= k["two_norm_reference"]
    vo, Vo, bo, _ = O.lanczos_eigenpairs(A, k["k"], start_vector(seed, 5), residual_row=row)
    v = vals.cpu().numpy()
    assert len(v) == len(vo) == 3
    assert np.allclose(v, vo, rtol=1e-10, atol=1e-12)
    assert bound == pytest.approx(bo, rel=1e-10)
    W = vecs.cpu().numpy()
    for j in range(3):
        s = np.sign(W[:, j] @ Vo[:, j])
        assert np.abs(W[:, j] - s * Vo[:, j]).max() <= 1e-8


@pytest.mark.parametrize("row", ["first", "last"])
@pytest.mark.parametrize("n,max_iter,kind", [(300, 25, "normal"), (700, 40, "normal_cluster")])
def test_eigenpairs_short_run_matches_oracle(handle, row, n, max_iter, kind):
    X, _, lam = O.spectral_input(n, 21, kind)
    seed = 41
    handle.set_seed(seed)
    handle.set_eigenpair_options(row)
    print(f"eigenpairs n={n} max_iter={max_iter} row={row}", flush=True)
    vals, vecs, bound = P.compute_eigenpairs(handle, to_dev(X), 10, max_iter=max_iter, ortho_tol=1e-3)
    handle.set_eigenpair_options("first")
    vo, Vo, bo, _ = O.lanczos_eigenpairs(X, 10, start_vector(seed, n), max_iter, 1e-3, residual_row=row)
    v = vals.cpu().numpy()
    assert len(v) == len(vo)
    assert np.allclose(v, vo, rtol=1e-10, atol=1e-12)
    assert bound == pytest.approx(bo, rel=1e-10)
    W = vecs.cpu().numpy()
    for j in range(len(v)):
        s = np.sign(W[:, j] @ Vo[:, j])
        assert np.abs(W[:, j] - s * Vo[:, j]).max() <= 1e-8


def test_eigenpairs_full_length_are_eigenpairs(handle):
    # max_iter = n (the default): orthogonality is lost, so compare with the
    # mathematics rather than with the oracle's (equally valid) iterates.
    n = 512
    X, _, lam = O.spectral_input(n, 23, "uniform")
    handle.set_seed(5)
    handle.set_eigenpair_options("last")
    print("eigenpairs full length n=512", flush=True)
    vals, vecs, bound = P.compute_eigenpairs(handle, to_dev(X), 20, ortho_tol=1e-5)
    handle.set_eigenpair_options("first")
    v = vals.cpu().numpy()
    W = vecs.cpu().numpy()
    assert len(v) >= 5
    assert np.all(np.diff(np.abs(v)) <= 1e-12)        # decreasing |theta|
    for j in range(len(v)):
        x = W[:, j]
        r = np.linalg.norm(X @ x - v[j] * x)
        assert r <= 1e-4 * max(1.0, np.linalg.norm(x))
        # an eigenvalue of A lies within ||r|| / ||x|| of theta (A symmetric)
        assert np.abs(lam - v[j]).min() <= r / np.linalg.norm(x) * (1 + 1e-6) + 1e-12
    G = W.T @ W
    off = G - np.diag(np.diag(G))
    assert np.abs(off).max() <= 1e-5
    assert bound >= np.abs(lam).max()


def test_eigenpairs_normalized_vectors(handle):
    n = 512
    X, _, lam = O.spectral_input(n, 23, "uniform")
    handle.set_seed(5)
    vals, raw, _ = P.compute_eigenpairs(handle, to_dev(X), 20, ortho_tol=1e-5)
    handle.set_eigenpair_options("first", normalize=True)
    vals2, unit, _ = P.compute_eigenpairs(handle, to_dev(X), 20, ortho_tol=1e-5)
    handle.set_eigenpair_options("first")
    assert torch.equal(vals, vals2)                     # same selection
    R, U = raw.cpu().numpy(), unit.cpu().numpy()
    assert np.allclose(np.linalg.norm(U, axis=0), 1.0, atol=1e-12)
    assert np.allclose(U * np.linalg.norm(R, axis=0), R, atol=1e-12)


def test_eigenpairs_picks_by_absolute_value(handle):
    vals, _, bound = P.compute_eigenpairs(handle, to_dev(np.diag([1.0, -2.0, 3.0, -4.0, 5.0])), 2)
    assert np.allclose(vals.cpu().numpy(), [5.0, -4.0], atol=1e-10)
    assert bound == pytest.approx(5.0 + 1e-3, abs=1e-9)


def test_eigenpairs_upper_bound_only(handle):
    X, _, lam = O.spectral_input(256, 25, "uniform_cluster")
    seed = 43
    handle.set_seed(seed)
    _, _, b = P.compute_eigenpairs(handle, to_dev(X), 0, upper_bound_only=True, max_iter=30, ortho_tol=1e-5)
    _, _, bo, _ = O.lanczos_eigenpairs(X, 0, start_vector(seed, 256), 30, 1e-5, upper_bound_only=True)
    assert b == pytest.approx(bo, rel=1e-10)


# ------------------------------------------------------------------ deflated projections
def test_fp32_deflate_kat_simple5(handle):
    k = KATS["simple5_deflate"]
    d = to_dev(np.diag(k["diagonal"]))
    P.composite_FP32_auto_scale_deflate(handle, d, k=k["k"])
    assert np.abs(from_dev(d) - np.diag(k["expected_diagonal"])).max() <= k["tol"]


def test_fp32_deflate_generate_and_project_1024(handle):
    # composite_FP32_test.hpp:292-334 (k = 40, tol 1e-10, ortho_tol 1e-5)
    A, Pm, _, _ = O.generate_and_project(1024, 10.0)
    seed = 51
    handle.set_seed(seed)
    print("fp32 deflate n=1024", flush=True)
    d = to_dev(A)
    P.composite_FP32_auto_scale_deflate(handle, d, k=40, tol=1e-10, ortho_tol=1e-5)
    Y = from_dev(d)
    err = np.linalg.norm(Y - Pm)
    assert err <= 1e-1
    Yo = O.composite_FP32_auto_scale_deflate(A, k=40, ortho_tol=1e-5, q0=start_vector(seed, 1024))
    e_gpu, e_or = O.rel_fro_error(Y, Pm), O.rel_fro_error(Yo, Pm)
    print(f"  rel err gpu {e_gpu:.3e} oracle {e_or:.3e}", flush=True)
    assert e_gpu <= 2 * e_or + 1e-6


def test_fp32_deflate_last_row_residuals_deflate_more(handle):
    A, Pm, _, _ = O.generate_and_project(512, 10.0)
    handle.set_seed(52)
    handle.set_eigenpair_options("last")
    d = to_dev(A)
    P.composite_FP32_auto_scale_deflate(handle, d, k=40, ortho_tol=1e-5)
    handle.set_eigenpair_options("first")
    assert O.rel_fro_error(from_dev(d), Pm) <= 1e-4


@pytest.mark.parametrize("n,k", [(100, 10), (1000, 100)])
def test_fp16_deflate_random(handle, n, k):
    # composite_FP16_test.hpp:67-138 (rand()/RAND_MAX entries, A + A^T, k = n/10)
    A = rand_symmetric(n, seed=3)
    seed = 61
    handle.set_seed(seed)
    print(f"fp16 deflate random n={n} k={k}", flush=True)
    d = to_dev(A)
    P.composite_FP16_auto_scale_deflate(handle, d, k=k)
    Y = from_dev(d)
    Pm = O.eig_FP64_psd(A)
    e = O.rel_fro_error(Y, Pm)
    print(f"  rel err {e:.3e}", flush=True)
    assert e <= 1e-3
    if n <= 100:
        Yo = O.composite_FP16_auto_scale_deflate(A, k=k, X0=start_block(seed, n, k), q0=start_vector(seed, n))
        assert e <= 1.5 * O.rel_fro_error(Yo, Pm) + 1e-5


def test_fp16_deflate_generate_and_project_1024(handle):
    # composite_FP16_test.hpp:19-65 (k = 100)
    A, Pm, _, _ = O.generate_and_project(1024, 10.0)
    handle.set_seed(62)
    print("fp16 deflate n=1024 k=100", flush=True)
    d = to_dev(A)
    P.composite_FP16_auto_scale_deflate(handle, d, k=100)
    e = O.rel_fro_error(from_dev(d), Pm)
    print(f"  rel err {e:.3e}", flush=True)
    assert e <= 5e-3


def test_eig_fp64_deflate_equals_evd(handle):
    A = rand_symmetric(400, seed=6)
    d = to_dev(A)
    P.eig_FP64_deflate(handle, d, k=20)
    assert O.rel_fro_error(from_dev(d), O.eig_FP64_psd(A)) <= 1e-10


def test_deflate_k0_is_plain_scaled_projection(handle):
    X, Q, lam = O.spectral_input(300, 27, "uniform")
    Pm = O.closed_form_projection(Q, lam)
    d = to_dev(X)
    P.composite_FP32_auto_scale_deflate(handle, d, k=0)
    assert O.rel_fro_error(from_dev(d), Pm) <= 1e-4
    d = to_dev(X)
    P.composite_FP16_auto_scale_deflate(handle, d, k=0)
    assert O.rel_fro_error(from_dev(d), Pm) <= 5e-3
