"""Pins the CPU oracle (oracle/psdp_oracle.py) before any GPU result is
compared with it: reference known-answer tests, values fixed by the
mathematics, independent library routines, and planted mistakes that must be
caught.  CPU only."""
import ctypes
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import psdp_oracle as O

GOLD = Path(__file__).resolve().parent / "golden"
KATS = json.loads((GOLD / "reference_kats.json").read_text())


# ----------------------------------------------------------------- generators
def test_glibc_rand_matches_libc():
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    expect = [libc.rand() for _ in range(5000)]
    got = O.GlibcRandom(1).take(5000).tolist()
    assert got == expect
    # the well-known first outputs after srand(1)
    assert got[:3] == [1804289383, 846930886, 1681692777]


def test_glibc_rand_chunked_matches_single_call():
    g = O.GlibcRandom(1)
    a = np.concatenate([g.take(7), g.take(100), g.take(1)])
    assert a.tolist() == O.GlibcRandom(1).take(108).tolist()


def test_generate_and_project_spectrum():
    A, P, Q, lam = O.generate_and_project(40, radius=10.0)
    assert np.allclose(A, A.T, atol=0)
    assert np.allclose(np.linalg.eigvalsh(A), np.sort(lam), atol=1e-12)
    assert np.allclose(Q.T @ Q, np.eye(40), atol=1e-13)
    assert np.allclose(P, O.closed_form_projection(Q, lam), atol=1e-12)


# ----------------------------------------------------------------- rounding
def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32),
                        np.array([1 + 2**-8, 1 + 3 * 2**-8, 1 + 2**-9, -1 - 2**-8, 0.0, 1e-40], np.float32)])
    t = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(O.round_bf16(x), t)


def test_f16_rounding_ties_to_even():
    x = np.array([1 + 2**-11, 1 + 3 * 2**-11, 2**-25, 3 * 2**-26], np.float32)
    assert O.round_f16(x).tolist() == [1.0, 1 + 2**-9, 0.0, 2**-24]


@pytest.mark.parametrize("fmt,bits", [("f16x3", 21), ("bf16x3", 15), ("bf16x9", 23)])
def test_split_reconstructs(fmt, bits):
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(10000) * 0.01).astype(np.float32)
    p = O.split_operand(x, fmt)
    if fmt == "f16x3":
        r = (p[0].astype(np.float64) + p[1].astype(np.float64)) / O.F16X3_SCALE
    else:
        r = sum(q.astype(np.float64) for q in p)
    # absolute error relative to the magnitude of the operand (tiny entries
    # fall into the binary16 subnormal range, where relative precision drops)
    assert np.max(np.abs(r - x)) / np.abs(x).max() < 2.0 ** -bits


def test_emulated_products():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((96, 96)).astype(np.float32) / 10
    Y = rng.standard_normal((96, 96)).astype(np.float32) / 10
    exact = X.astype(np.float64) @ Y.astype(np.float64)
    err = lambda f: np.abs(O.emulated_product(X, Y, f) - exact).max() / np.abs(exact).max()
    # float32 accumulation bounds everything at ~1e-7..1e-6; the rounding of
    # the operands sets the rest (2^-11 f16, 2^-8 bf16, 2^-16 bf16x3).
    assert err("f32") < 1e-6
    assert err("bf16x9") < 1e-6
    assert err("f16x3") < 1e-6
    assert 1e-6 < err("bf16x3") < 2e-5
    assert 1e-4 < err("f16") < 1e-3
    assert 1e-3 < err("bf16") < 1e-2


# ----------------------------------------------------------------- chain
def test_transfer_function_max_error():
    # max |g - ReLU| on [-1, 1]: 4.48e-6 (FP32 table), 4.65e-5 (FP16 table)
    x = np.linspace(-1, 1, 200001)
    relu = np.maximum(x, 0)
    e32 = np.abs(O.transfer_function(x, O.fp32_chain_coefficients()) - relu).max()
    e16 = np.abs(O.transfer_function(x, O.fp16_chain_coefficients()) - relu).max()
    assert e32 == pytest.approx(4.48e-6, rel=0.01)
    assert e16 == pytest.approx(4.65e-5, rel=0.01)


def test_transfer_function_shape():
    c = O.fp32_chain_coefficients()
    P = lambda x: O.polynomial_chain_scalar(np.array(x), c)
    xs = np.linspace(1e-3, 1.0, 1000)
    assert np.abs(P(xs) - 1).max() < 1e-5
    assert np.allclose(P(-xs), -P(xs), atol=0)  # odd
    assert P(1e-4) == pytest.approx(1.0013, abs=2e-3)
    assert P(3e-5) == pytest.approx(0.775, abs=5e-3)
    assert P(1e-5) == pytest.approx(0.31, abs=1e-2)


@pytest.mark.parametrize("table,ref", [(O.COEFF_FP32_TABLE, 4.48e-6), (O.COEFF_FP16_TABLE, 4.65e-5)])
def test_coefficient_scaling_planted_mistake(table, ref):
    # Planted mistake: coefficients without the 1.001 / 1.01 scale factor
    # miss the max-error value above by more than its tolerance.
    bad = O.step_coefficients(table, 1.0, 0)
    x = np.linspace(-1, 1, 200001)
    e_bad = np.abs(O.transfer_function(x, bad) - np.maximum(x, 0)).max()
    assert abs(e_bad / ref - 1) > 0.02


def test_chain_matches_closed_form():
    d = np.load(GOLD / "chain_small.npz")
    for n in (64, 128):
        X, Q, lam = d[f"X{n}"], d[f"Q{n}"], d[f"lam{n}"]
        for table, coeffs in (("fp32", O.fp32_chain_coefficients()), ("fp16", O.fp16_chain_coefficients())):
            Y = O.composite_exact(X, table)
            G = (Q * O.transfer_function(lam, coeffs)) @ Q.T
            assert np.abs(Y - G).max() < 1e-12
            assert np.abs(Y - d[f"chain_{table}_table_f64_{n}"]).max() < 1e-13


def test_chain_planted_mistake_caught():
    d = np.load(GOLD / "chain_small.npz")
    X, Q, lam = d["X64"], d["Q64"], d["lam64"]
    c = O.fp32_chain_coefficients()
    wrong = [(a, b, 0.0) for a, b, _ in c]  # dropped c x^5 term
    with np.errstate(all="ignore"):
        Y = O.composite_chain(X, wrong, ["f64"] * len(c), "f64")
    G = (Q * O.transfer_function(lam, c)) @ Q.T
    assert not np.all(np.abs(Y - G) <= 1e-3)


def test_kat_deterministic_3x3():
    k = KATS["composite_fp32_deterministic"]
    X = np.array(k["input_colmajor"]).reshape(3, 3, order="F")
    Y = O.composite_FP32(X)
    E = np.array(k["expected_colmajor"]).reshape(3, 3, order="F")
    assert np.abs(Y - E).max() <= k["tol"]


@pytest.mark.parametrize("method", ["fp32", "fp16", "fp32_emulated", "exact"])
def test_kat_simple5(method):
    # The reference checks this KAT at 1e-3 with an FP32 chain; binary16
    # operands carry 2^-11 relative error on entries up to 5, so the FP16
    # chain is held to 1e-3 of the largest entry instead.
    k = KATS["simple5"]
    X = np.diag(k["diagonal"])
    Y = O.composite_auto_scale(X, method, q0=np.ones(5))
    tol = k["tol"] * (5.0 if method == "fp16" else 1.0)
    assert np.abs(Y - np.diag(k["expected_diagonal"])).max() <= tol


def test_fp32_chain_error_level():
    X, Q, lam = O.spectral_input(256, 1, "uniform")
    P = O.closed_form_projection(Q, lam)
    _, up = O.lanczos_two_norm(X)
    err = O.rel_fro_error(O.composite_FP32(X / up) * up, P)
    assert 1e-6 < err < 2e-5


def test_fp16_chain_error_level_clustered():
    # The reference FP16 chain misses 1e-3 on the normal_cluster spectrum.
    X, Q, lam = O.spectral_input(512, 2, "normal_cluster")
    P = O.closed_form_projection(Q, lam)
    _, up = O.lanczos_two_norm(X)
    err = O.rel_fro_error(O.composite_FP16(X / up) * up, P)
    assert 1.5e-3 < err < 5e-3


# ----------------------------------------------------------------- Lanczos
def test_lanczos_kat_5x5():
    k = KATS["lanczos_5x5"]
    A = np.array(k["input_colmajor"]).reshape(5, 5, order="F")
    assert k["two_norm"] == pytest.approx(7.089012550562305, abs=1e-12)
    lo, up = O.lanczos_two_norm(A, k["max_iter"], k["tol"], q0=np.ones(5))
    assert lo <= k["two_norm"] + 1e-9 <= up + 1e-9
    # n = 5: the fifth beta vanishes, the break rule stops at nb = 4, and the
    # 4x4 Krylov space already pins the top Ritz value to ~1e-7.
    assert lo == pytest.approx(k["two_norm"], abs=1e-6)


@pytest.mark.parametrize("lam", [-10.0, 3.0, 0.5, -0.7])
@pytest.mark.parametrize("iters", [2, 3, 4])
def test_lanczos_matches_rayleigh_ritz(lam, iters):
    # Independent route to the same bound: an orthonormal basis of the Krylov
    # space span{q, B q, ..., B^(k-1) q} of B = A^T A by QR, the top
    # Rayleigh-Ritz pair (theta, y) of B on it, and the reference's residual
    # r = A y - theta y (src/lanczos.cu:255 copies ry into w1).
    rng = np.random.default_rng(7)
    d = np.concatenate([[lam], rng.uniform(-abs(lam) / 2, abs(lam) / 2, 39)])
    Qr, _ = np.linalg.qr(rng.standard_normal((40, 40)))
    A = (Qr * d) @ Qr.T
    q0 = np.ones(40)
    B = A.T @ A
    K = np.stack([np.linalg.matrix_power(B, j) @ q0 for j in range(iters)], axis=1)
    Qk, _ = np.linalg.qr(K)
    w, U = np.linalg.eigh(Qk.T @ B @ Qk)
    y = Qk @ U[:, -1]
    up_rr = np.sqrt(w[-1] + np.linalg.norm(A @ y - w[-1] * y))
    lo, up = O.lanczos_two_norm(A, iters, 0.0, q0=q0)
    assert lo == pytest.approx(np.sqrt(w[-1]), rel=1e-9)
    assert up == pytest.approx(up_rr, rel=1e-7)
    # planted mistake: the textbook residual B y - theta y gives another bound
    up_textbook = np.sqrt(w[-1] + np.linalg.norm(B @ y - w[-1] * y))
    assert abs(up_textbook - up) > 1e-3 * up


@pytest.mark.parametrize("lam", [-10.0, 3.0, 0.5])
@pytest.mark.parametrize("iters", [2, 3, 4])
def test_lanczos_tight_matches_rayleigh_ritz(lam, iters):
    # The tight mode's residual is that of the Rayleigh-Ritz pair of
    # B = A^T A itself; the same independent Krylov/QR route gives it.
    rng = np.random.default_rng(7)
    d = np.concatenate([[lam], rng.uniform(-abs(lam) / 2, abs(lam) / 2, 39)])
    Qr, _ = np.linalg.qr(rng.standard_normal((40, 40)))
    A = (Qr * d) @ Qr.T
    q0 = np.ones(40)
    B = A.T @ A
    K = np.stack([np.linalg.matrix_power(B, j) @ q0 for j in range(iters)], axis=1)
    Qk, _ = np.linalg.qr(K)
    w, U = np.linalg.eigh(Qk.T @ B @ Qk)
    y = Qk @ U[:, -1]
    up_tb = np.sqrt(w[-1] + np.linalg.norm(B @ y - w[-1] * y))
    lo, up = O.lanczos_two_norm(A, iters, 0.0, q0=q0, mode="tight")
    assert lo == pytest.approx(np.sqrt(w[-1]), rel=1e-9)
    assert up == pytest.approx(up_tb, rel=1e-7)


def test_lanczos_tight_complete_krylov_space():
    # A Krylov space that holds every eigenvector of B (n = 5, 50 iterations):
    # theta = ||A||_2^2 and the Ritz residual vanishes, so the tight bound is
    # ||A||_2 itself; the reference residual A y - theta y of the same pair is
    # |lambda - lambda^2| and leaves up far above.
    k = KATS["lanczos_5x5"]
    A = np.array(k["input_colmajor"]).reshape(5, 5, order="F")
    lo, up = O.lanczos_two_norm(A, 50, 0.0, q0=np.ones(5), mode="tight")
    assert up == pytest.approx(k["two_norm"], rel=1e-7)
    assert lo == pytest.approx(k["two_norm"], rel=1e-9)
    _, up_ref = O.lanczos_two_norm(A, 50, 0.0, q0=np.ones(5))
    nrm = k["two_norm"]
    assert up_ref == pytest.approx(np.sqrt(nrm ** 2 + abs(nrm - nrm ** 2)), rel=1e-6)


@pytest.mark.parametrize("kind", ["uniform", "normal", "uniform_cluster"])
def test_lanczos_tight_brackets_norm(kind):
    X, Q, lam = O.spectral_input(300, 11, kind)
    nrm = np.abs(lam).max()
    lo, up = O.lanczos_two_norm(X, 20, 1e-10, mode="tight")
    _, up_ref = O.lanczos_two_norm(X, 20, 1e-10)
    assert lo <= nrm * (1 + 1e-12) <= up
    assert up <= up_ref


def test_lanczos_zero_matrix():
    assert O.lanczos_two_norm(np.zeros((6, 6))) == (0.0, 1.0)


@pytest.mark.parametrize("seed", range(6))
def test_lanczos_brackets_norm(seed):
    rng = np.random.default_rng(seed)
    n = 200
    G = rng.standard_normal((n, n))
    A = (G + G.T) / 2
    if seed % 2:
        A -= 3 * np.abs(np.linalg.eigvalsh(A)).max() * np.outer(np.ones(n), np.ones(n)) / n
    nrm = np.abs(np.linalg.eigvalsh(A)).max()
    lo, up = O.lanczos_two_norm(A, 20, 1e-10)
    assert lo <= nrm * (1 + 1e-12)
    assert up >= nrm


# ----------------------------------------------------------------- EVD
def test_eig_fp64_kats():
    k = KATS["simple5"]
    Y, w = O.eig_FP64_psd(np.diag(k["diagonal"]), return_eigenvalues=True)
    assert np.abs(Y - np.diag(k["expected_diagonal"])).max() < 1e-14
    assert np.allclose(w, np.sort(k["diagonal"]))
    k3 = KATS["composite_fp32_deterministic"]
    X = np.array(k3["input_colmajor"]).reshape(3, 3, order="F")
    assert np.abs(O.eig_FP64_psd(X) - X).max() < 1e-14


def test_eig_fp64_closed_form():
    X, Q, lam = O.spectral_input(150, 5, "normal")
    assert np.abs(O.eig_FP64_psd(X) - O.closed_form_projection(Q, lam)).max() < 1e-12


def test_eig_fp32_close():
    X, Q, lam = O.spectral_input(120, 6, "normal")
    assert O.rel_fro_error(O.eig_FP32_psd(X), O.closed_form_projection(Q, lam)) < 1e-5


def test_kat_lobpcg_eigenvalues():
    k = KATS["lobpcg_6x6"]
    ev = sorted(k["eigenvalues"], reverse=True)[:2]
    assert np.allclose(ev, k["expected_top2_reference"], atol=k["tol"])


# ----------------------------------------------------------------- symmetrization
def test_average_mode_tracks_reference_symmetrize():
    # Averaging every product with its transpose keeps the reference's error
    # (which symmetrizes A once per step by averaging); mirroring the upper
    # triangle of plain-FP16 products roughly doubles it.
    X, Q, lam = O.spectral_input(384, 2, "normal_cluster")
    Pt = O.closed_form_projection(Q, lam)
    _, up = O.lanczos_two_norm(X)
    Xs = (X / up).astype(np.float32)
    c = O.fp16_chain_coefficients()
    err = {m: O.rel_fro_error(O.composite_chain(Xs, c, ["f16"] * 7, "f16", sym_mode=m) * up, Pt)
           for m in ("reference", "average", "mirror")}
    assert err["average"] == pytest.approx(err["reference"], rel=0.15)
    assert err["mirror"] > 1.5 * err["reference"]


def test_mirror_and_average_products():
    rng = np.random.default_rng(3)
    G = rng.standard_normal((40, 40))
    X = ((G + G.T) / 8).astype(np.float32)
    Y = (X.astype(np.float64) @ X).astype(np.float32)
    C = O.emulated_product(X, Y, "f16")
    M = O.sym_product(X, Y, "f16", "mirror")
    A = O.sym_product(X, Y, "f16", "average")
    iu = np.triu_indices(40)
    assert np.array_equal(M[iu], C[iu]) and np.array_equal(M, M.T)
    assert np.allclose(A, (C + C.T) / 2) and np.array_equal(A, A.T)


@pytest.mark.parametrize("a", [-3.0, 2.0, 0.25, 1e-3])
def test_lanczos_one_by_one(a):
    # 1x1: q is an eigenvector, beta_0 = 0; the exact Ritz value a^2 is kept
    # and up = sqrt(a^2 + |a - a^2|) >= |a|.  The reference's own rule would
    # report (0, 1) here (it drops the iteration), which the chain then
    # applies to an unscaled |a| > 1.
    lo, up = O.lanczos_two_norm(np.array([[a]]), 20, 1e-10, q0=np.ones(1))
    assert lo == pytest.approx(abs(a), rel=1e-15)
    assert up == pytest.approx(np.sqrt(a * a + abs(a - a * a)), rel=1e-15)
    assert O.lanczos_two_norm(np.array([[a]]), 20, 1e-10, q0=np.ones(1), keep_exact=False) == (0.0, 1.0)


def test_lanczos_keeps_krylov_completing_iteration():
    # 3 x 3 input (rng 148) and the start vector of seed 9: the reference's
    # convergence rule (src/lanczos.cu:145) drops the third iteration, which
    # completes the Krylov space, and its bound then undercuts ||X||_2; the
    # library keeps that iteration (keep_exact), so theta is ||X||_2^2.
    from _streams import start_vector
    rng = np.random.default_rng(148)
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    lam = rng.uniform(-5.0, 5.0, 3)
    X = (Q * lam) @ Q.T
    X = 0.5 * (X + X.T)
    nrm = np.linalg.norm(X, 2)
    q0 = start_vector(9, 3)
    lo0, up0 = O.lanczos_two_norm(X, 20, 1e-10, q0, keep_exact=False)
    lo1, up1, st = O.lanczos_two_norm(X, 20, 1e-10, q0, return_state=True)
    assert up0 < nrm
    assert st["nb"] == 3
    assert lo1 == pytest.approx(nrm, rel=1e-12) and up1 >= nrm
    # a larger matrix is unaffected (the rule only applies at k = n - 1)
    Y = np.diag(np.linspace(-2.0, 3.0, 40))
    assert O.lanczos_two_norm(Y, 20, 1e-10, start_vector(9, 40)) == \
        O.lanczos_two_norm(Y, 20, 1e-10, start_vector(9, 40), keep_exact=False)


def test_lanczos_invariant_start_vector():
    # q0 = e_0 spans an eigenspace of diag(d): beta_0 = 0 at once
    d = np.array([-4.0, 1.0, 2.0, 3.0])
    lo, up = O.lanczos_two_norm(np.diag(d), 20, 1e-10, q0=np.eye(4)[0])
    assert lo == pytest.approx(4.0) and up == pytest.approx(np.sqrt(16 + 20))


# ------------------------------------------------------------------ e4m3 grid
def test_round_e4m3_known_answers():
    v = np.array([0.0, 448.0, 460.0, 1e6, -1e6, 2.0 ** -6, 2.0 ** -9, 2.0 ** -10, 1.5 * 2.0 ** -10,
                  1.0 + 1.0 / 16, 1.0 + 3.0 / 16, 17.0, 19.0], dtype=np.float32)
    want = np.array([0.0, 448.0, 448.0, 448.0, -448.0, 2.0 ** -6, 2.0 ** -9, 0.0, 2.0 ** -9,
                     1.0, 1.25, 16.0, 20.0], dtype=np.float32)
    # 2^-10 is a tie between 0 and 2^-9 (even: 0); 1 + 1/16 ties to 1, 1 + 3/16 to 1.25;
    # 17 = 16 + 1 ties to 16 (quantum 2 above 16), 19 to 20
    assert np.array_equal(O.round_e4m3(v), want)


def test_round_e4m3_is_the_code_grid():
    torch = pytest.importorskip("torch")
    codes = torch.arange(256, dtype=torch.uint8).view(torch.float8_e4m3fn).float().numpy()
    fin = codes[np.isfinite(codes)]
    assert len(fin) == 254 and fin.max() == 448.0
    assert np.array_equal(O.round_e4m3(fin), fin)
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(50000) * rng.choice([1e-3, 1.0, 30.0, 300.0], 50000)).astype(np.float32)
    x = x[np.abs(x) < 440]
    assert np.array_equal(O.round_e4m3(x), torch.from_numpy(x).to(torch.float8_e4m3fn).float().numpy())
    # planted mistake: truncation instead of round-to-nearest fails
    trunc = np.sign(x) * np.floor(np.abs(x) * 8) / 8
    assert not np.array_equal(O.round_e4m3(x), trunc)


def test_e4m3_product_error_band():
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, (64, 64)).astype(np.float32) / 8
    exact = X.astype(np.float64) @ X.astype(np.float64)
    e = np.linalg.norm(O.emulated_product(X, X, "e4m3") - exact) / np.linalg.norm(exact)
    # 3 mantissa bits: relative operand error <= 2^-4, product error a few 1e-2
    assert 5e-3 < e < 6e-2, e


def test_mx_e4m3_known_answers():
    # block max 1.0 (exponent field 127) -> scale 2^-7: scaled values up to 128,
    # so values on the e4m3 grid times 2^-7 come back unchanged
    blk = np.array([1.0, 0.5, -0.25, 0.75, 2.0 ** -9, 0.0] + [0.125] * 26, dtype=np.float32)
    assert np.array_equal(O.mx_e4m3(blk[None, :], 1)[0], blk)
    # max 1.99 keeps scale 2^-7 (scaled max 254.7 -> 256, no saturation at 448)
    b2 = np.zeros(32, np.float32)
    b2[0], b2[1] = 1.99, 1.0 + 1.0 / 16  # 1 + 1/16 -> 136 scaled: quantum 16 -> 128 (tie to even) -> 1.0
    got = O.mx_e4m3(b2[None, :], 1)[0]
    assert got[0] == 2.0 and got[1] == 1.0
    # an all-zero block and a block of one tiny value
    assert not O.mx_e4m3(np.zeros((3, 32), np.float32), 1).any()
    t = np.zeros(32, np.float32)
    t[5] = 3e-30
    g = O.mx_e4m3(t[None, :], 1)[0]
    assert g[5] == pytest.approx(3e-30, rel=2 ** -4) and not np.delete(g, 5).any()


def test_mx_e4m3_brute_force():
    torch = pytest.importorskip("torch")
    codes = torch.arange(256, dtype=torch.uint8).view(torch.float8_e4m3fn).float().numpy().astype(np.float64)
    grid = np.unique(codes[np.isfinite(codes)])
    rng = np.random.default_rng(11)
    x = (rng.standard_normal((40, 96)) * np.exp2(rng.integers(-20, 8, (40, 1)))).astype(np.float32)
    x[3, 40:70] = 0.0
    got = O.mx_e4m3(x, 1).astype(np.float64)
    for r in range(x.shape[0]):
        for k0 in range(0, 96, 32):
            blk = x[r, k0:k0 + 32].astype(np.float64)
            m = np.abs(blk).max()
            # scale 2^(floor(log2 m) - 7) from frexp (m = f 2^e, f in [0.5, 1))
            sc = 1.0 if m == 0 else 2.0 ** (np.frexp(m)[1] - 1 - 7)
            assert m / sc < 256
            for k in range(32):
                v = blk[k] / sc
                cand = grid[np.abs(grid - v) == np.abs(grid - v).min()]
                if len(cand) > 1:  # tie: the even code (last mantissa bit 0)
                    cand = [c for c in cand if (torch.tensor(c, dtype=torch.float32).to(torch.float8_e4m3fn)
                                                .view(torch.uint8).item() & 1) == 0]
                assert got[r, k0 + k] == cand[0] * sc
    # blocks along axis 0 are blocks along axis 1 of the transpose; a ragged end is zero padded
    assert np.array_equal(O.mx_e4m3(x, 0), O.mx_e4m3(x.T, 1).T)
    xr = x[:, :70]
    assert np.array_equal(O.mx_e4m3(xr, 1), O.mx_e4m3(np.pad(xr, ((0, 0), (0, 26))), 1)[:, :70])


def test_f16mx_product_error_band():
    # hi.hi exact in float32 plus cross terms from e4m3 copies (relative
    # error 2^-4 of terms 2^-11 below hi.hi): ~1e-5, far below binary16
    # (~3e-4) and above FP16x3 (~3e-7).  Planted mistakes -- cross terms
    # dropped, paired hi8.hi8 / lo8.lo8, or blocked along the wrong axis of
    # a non-symmetric operand -- leave the band.
    X, _, _ = O.spectral_input(256, 5, "uniform")
    Y, _, lam = O.spectral_input(256, 9, "normal")
    Y = Y / np.abs(lam).max()
    X32, Y32 = X.astype(np.float32), Y.astype(np.float32)
    exact = X32.astype(np.float64) @ Y32.astype(np.float64)
    rel = lambda c: np.linalg.norm(c - exact) / np.linalg.norm(exact)
    e_mx = rel(O.emulated_product(X32, Y32, "f16mx"))
    e16, e3 = rel(O.emulated_product(X32, Y32, "f16")), rel(O.emulated_product(X32, Y32, "f16x3"))
    assert 3e-6 < e_mx < 3e-5 and e_mx < e16 / 10 and e_mx > 10 * e3, (e_mx, e16, e3)
    px, py = O.mx_split(X32, 1), O.mx_split(Y32, 0)
    s16 = np.float32(1.0 / 65536)
    mm = lambda a, b: (a @ b).astype(np.float32)
    assert rel(mm(px[0], py[0]) * s16) > e16 / 4  # no cross terms
    assert rel((mm(px[1], py[1]) + mm(px[0], py[0])) * s16) > 0.5  # hi8.hi8 instead of hi8.lo8
    Z = np.triu(Y32) * np.float32(3)  # non-symmetric: the k axis of y is axis 0
    pz_wrong = O.mx_split(Z, 1)
    ez = np.linalg.norm(O.emulated_product(X32, Z, "f16mx") - X32.astype(np.float64) @ Z) / np.linalg.norm(
        X32.astype(np.float64) @ Z)
    ew = np.linalg.norm((mm(px[1], pz_wrong[2]) + mm(px[2], pz_wrong[1]) + mm(px[0], pz_wrong[0])) * s16
                        - X32.astype(np.float64) @ Z) / np.linalg.norm(X32.astype(np.float64) @ Z)
    assert ez < 3e-5 and ew > ez


def test_mixed_mx_chain_error():
    # the mixed_mx schedule keeps the 1e-3 bar, a little above mixed
    X, Q, lam = O.spectral_input(512, 3, "uniform")
    Pt = O.closed_form_projection(Q, lam)
    c = O.fp16_chain_coefficients()
    e_mx = O.rel_fro_error(O.composite_chain(X, c, ["f16mx"] * 4 + ["f16"] * 3, "f16", sym_mode="mirror"), Pt)
    e_m = O.rel_fro_error(O.composite_chain(X, c, ["f16x3"] * 4 + ["f16"] * 3, "f16", sym_mode="mirror"), Pt)
    assert e_mx <= 1e-3 and e_m <= e_mx <= 1.3 * e_m, (e_mx, e_m)


@pytest.mark.parametrize("kind", ["normal", "uniform"])
def test_eig_fp64_staged_matches_eigh(kind):
    # The timed CPU baseline (dsytrd + dstedc + dormtr, the stages of dsyevd)
    # equals the single-call dsyevd route and the closed form.
    X, Q, lam = O.spectral_input(257, 8, kind)
    Y, w, stages = O.eig_FP64_psd_staged(X)
    Y2, w2 = O.eig_FP64_psd(X, return_eigenvalues=True)
    assert set(stages) == {"dsytrd", "dstedc", "dormtr", "clamp_gemm"}
    assert np.abs(Y - Y2).max() < 1e-12
    assert np.abs(w - w2).max() < 1e-12
    assert O.rel_fro_error(Y, O.closed_form_projection(Q, lam)) < 1e-12
    k = KATS["simple5"]
    Y5, w5, _ = O.eig_FP64_psd_staged(np.diag(k["diagonal"]))
    assert np.abs(Y5 - np.diag(k["expected_diagonal"])).max() < 1e-14
    assert np.allclose(w5, np.sort(k["diagonal"]))


def test_f16x3_value_bound_and_exact_grid():
    # A^3 of a binary16x3 step is read back from its pieces: (hi + lo) 2^-8
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, 200000).astype(np.float32)
    v = O.f16x3_value(x)
    # relative 2^-22 from the 11 + 11 bits, absolute 2^-33 where lo falls into
    # the binary16 subnormals (quantum 2^-24, scaled back by 2^-8)
    assert np.all(np.abs(v.astype(np.float64) - x) <= 2.0 ** -22 * np.abs(x) + 2.0 ** -33)
    # values already on the grid come back exactly, and a dropped lo piece
    # (a plausible mistake) is caught by the bound
    hi, lo = O.split_operand(x, "f16x3")
    g = ((hi + lo) / np.float32(256.0)).astype(np.float32)
    assert np.array_equal(O.f16x3_value(g), g)
    assert np.any(np.abs(hi / np.float32(256.0) - x) > 2.0 ** -22 * np.abs(x))
