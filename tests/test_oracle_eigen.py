"""Pins the oracle's eigenpair and deflation restatements (CPU only):
reference KATs (lanczos_test.hpp, lobpcg_test.hpp, composite_FP32_test.hpp,
composite_FP16_test.hpp) and identities fixed by the mathematics."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import psdp_oracle as O

KATS = json.loads((Path(__file__).resolve().parent / "golden" / "reference_kats.json").read_text())


def rand_symmetric(n, seed=1):
    """lobpcg_test.hpp:78-89: rand()/RAND_MAX entries (filled in memory
    order), then A + A^T."""
    g = O.GlibcRandom(seed)
    M = g.take(n * n).astype(np.float64).reshape(n, n, order="F") / O.GlibcRandom.RAND_MAX
    return M + M.T


# ------------------------------------------------------------------ LOBPCG
def test_lobpcg_kat_6x6():
    k = KATS["lobpcg_6x6"]
    A = np.array(k["input_colmajor"]).reshape(6, 6, order="F")
    V, lam, it, rn = O.lobpcg(A, 2, np.random.default_rng(0).random((6, 2)), 100, 1e-8)
    assert np.allclose(lam, k["expected_top2_reference"], atol=k["tol"])
    assert np.allclose(lam, sorted(k["eigenvalues"])[::-1][:2], atol=1e-10)
    assert rn < 1e-8
    assert np.allclose(V.T @ V, np.eye(2), atol=1e-12)


@pytest.mark.parametrize("n,m", [(100, 10), (60, 20)])
def test_lobpcg_top_m_match_eigh(n, m):
    A = rand_symmetric(n)
    V, lam, it, rn = O.lobpcg(A, m, np.random.default_rng(3).random((n, m)), 100, 1e-8)
    ev = np.linalg.eigvalsh(A)[::-1][:m]
    assert np.abs(lam - ev).max() <= 1e-8          # lobpcg_test.hpp:110 bar
    assert np.all(np.diff(lam) <= 0)               # descending
    assert np.linalg.norm(A @ V - V * lam) <= 1e-7 * max(1.0, np.abs(ev).max())


def test_lobpcg_sign_gives_smallest():
    A = rand_symmetric(90, seed=2)
    _, mu, _, _ = O.lobpcg(A, 5, np.random.default_rng(4).random((90, 5)), 200, 1e-9, sign=-1.0)
    assert np.allclose(-mu, np.linalg.eigvalsh(A)[:5], atol=1e-8)


def test_lobpcg_warm_start_from_converged_block_stops_at_once():
    A = rand_symmetric(80, seed=3)
    V, lam, _, _ = O.lobpcg(A, 4, np.random.default_rng(5).random((80, 4)), 200, 1e-10)
    V2, lam2, it, rn = O.lobpcg(A, 4, V, 200, 1e-8, warm=True)
    assert it == 0 and rn < 1e-8
    assert np.allclose(lam2, lam, atol=1e-12)


# ------------------------------------------------------------------ Lanczos eigenpairs
def test_eigenpairs_kat_5x5():
    k = KATS["eigenpairs_5x5"]
    A = np.array(k["input_colmajor"]).reshape(5, 5, order="F")
    vals, V, bound, st = O.lanczos_eigenpairs(A, k["k"], np.linspace(1.0, 2.0, 5))
    assert bound >= k["two_norm_reference"]
    ev = np.array(k["eigenvalues"])
    top = ev[np.argsort(-np.abs(ev))][:3]
    assert len(vals) == 3 and np.allclose(vals, top, atol=1e-10)
    for j in range(3):
        x = V[:, j]
        assert np.linalg.norm(A @ x - vals[j] * x) <= 1e-9


def test_eigenpairs_picks_by_absolute_value():
    A = np.diag([1.0, -2.0, 3.0, -4.0, 5.0])
    vals, _, bound, _ = O.lanczos_eigenpairs(A, 2, np.linspace(1.0, 2.0, 5))
    assert np.allclose(vals, [5.0, -4.0], atol=1e-10)
    assert bound == pytest.approx(5.0 + 1e-3, abs=1e-9)


@pytest.mark.parametrize("row", ["first", "last"])
def test_eigenpair_residual_identities(row):
    # Before orthogonality is lost (short run), Q is orthonormal and
    #   last-row residual  = ||A x - theta x||   (x = Q z, unit)
    #   first-row residual = |beta_last (q0 . x)|
    rng = np.random.default_rng(7)
    n = 200
    X, _, lam = O.spectral_input(n, 7, "normal")
    q0 = rng.random(n)
    _, _, _, st = O.lanczos_eigenpairs(X, 0, q0, max_iter=12, upper_bound_only=True, residual_row=row)
    Q, Z, th = st["Q"], st["Z"], st["theta"]
    assert np.abs(Q.T @ Q - np.eye(Q.shape[1])).max() < 1e-10
    for j in range(st["nb"]):
        x = Q @ Z[:, j]
        if row == "last":
            expect = np.linalg.norm(X @ x - th[j] * x)
        else:
            expect = abs(st["beta"][-1] * (Q[:, 0] @ x))
        assert st["res"][j] == pytest.approx(expect, rel=1e-6, abs=1e-12)


def test_eigenpairs_full_length_small_matrix_is_exact():
    X, _, lam = O.spectral_input(40, 9, "uniform")
    vals, V, bound, st = O.lanczos_eigenpairs(X, 6, np.ones(40), residual_row="last", ortho_tol=1e-6)
    top = lam[np.argsort(-np.abs(lam))][:len(vals)]
    assert len(vals) >= 3
    assert np.allclose(vals, top, atol=1e-8)
    assert bound >= np.abs(lam).max()


# ------------------------------------------------------------------ deflation
def test_fp32_deflate_kat_simple5():
    k = KATS["simple5_deflate"]
    Y = O.composite_FP32_auto_scale_deflate(np.diag(k["diagonal"]), k=k["k"], q0=np.linspace(1.0, 2.0, 5))
    assert np.abs(Y - np.diag(k["expected_diagonal"])).max() <= k["tol"]


def test_fp32_deflate_needs_unit_ritz_vectors():
    # The reference's deflation (unnormalised x = Q z from a basis that lost
    # orthogonality, src/lanczos.cu:478-481) misses its own bar
    # (composite_FP32_test.hpp:329, ||Y - Pi||_F <= 0.1) at n = 1024; with the
    # kept vectors normalised (same selection) it reaches the FP32 chain's
    # accuracy.
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parent))
    from _streams import start_vector
    A, Pm, _, _ = O.generate_and_project(1024, 10.0)
    q0 = start_vector(1, 1024)
    raw = O.composite_FP32_auto_scale_deflate(A, k=40, ortho_tol=1e-5, q0=q0, normalize=False)
    unit = O.composite_FP32_auto_scale_deflate(A, k=40, ortho_tol=1e-5, q0=q0, normalize=True)
    assert np.linalg.norm(raw - Pm) > 0.1
    assert np.linalg.norm(unit - Pm) <= 1e-2
    assert O.rel_fro_error(unit, Pm) <= 1e-4


def test_fp32_deflate_generate_and_project():
    A, Pm, _, _ = O.generate_and_project(256, 10.0)
    Y = O.composite_FP32_auto_scale_deflate(A, k=40, ortho_tol=1e-5, q0=np.ones(256))
    assert np.linalg.norm(Y - Pm) <= 1e-1        # composite_FP32_test.hpp:329 bar
    assert O.rel_fro_error(Y, Pm) <= 1e-4        # the FP32 chain's own accuracy


def test_fp16_deflate_random_symmetric():
    A = rand_symmetric(100)
    Y = O.composite_FP16_auto_scale_deflate(A, k=10, X0=np.random.default_rng(1).random((100, 10)),
                                            q0=np.ones(100))
    assert O.rel_fro_error(Y, O.eig_FP64_psd(A)) <= 1e-3   # composite_FP16_test.hpp:129


def test_fp16_deflate_generate_and_project():
    A, Pm, _, _ = O.generate_and_project(256, 10.0)
    Y = O.composite_FP16_auto_scale_deflate(A, k=20, X0=np.random.default_rng(1).random((256, 20)),
                                            q0=np.ones(256))
    assert O.rel_fro_error(Y, Pm) <= 5e-3        # composite_FP16_test.hpp:57


def test_eig_deflate_equals_evd_projection():
    A = rand_symmetric(100, seed=4)
    Y = O.eig_FP64_deflate(A, k=10, X0=np.random.default_rng(2).random((100, 10)))
    assert O.rel_fro_error(Y, O.eig_FP64_psd(A)) <= 1e-10
