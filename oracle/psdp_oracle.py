"""CPU oracle for the PSD-projection hot path.  TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker (or, for the CPU
baseline, as the timed CPU port of the reference's ``eig_FP64_psd``).  The
library in ``psd_projection_amd`` never imports it and has no CPU fallback.

Everything here restates the reference algorithm in NumPy (float64 unless a
rounding format is emulated on purpose).  Each function cites the reference
file:line it follows.  How each function is pinned (golden vectors, closed
forms, invariants) is listed in PINNING at the bottom; tests/test_oracle.py
holds the checks.

Reference files (ComputationalRobotics/psd_projection):
  src/composite_FP32.cu, src/composite_FP16.cu, src/composite_FP32_emulated.cu,
  src/lanczos.cu, src/lobpcg.cu, src/eig_FP64_psd.cu, src/eig_FP32_psd.cu,
  src/utils.cu, test/test_utils.hpp
"""
from __future__ import annotations

import numpy as np

# ----------------------------------------------------------------------------
# Coefficient tables (data, kept identical to the reference on purpose: the
# method is defined by these numbers).
# ----------------------------------------------------------------------------

# src/composite_FP32.cu:73-84 ("10 polar express refined"); also used by
# src/composite_FP32_emulated.cu:78-89.  Stored by the reference as float.
COEFF_FP32_TABLE = (
    (8.3119043343, -23.0739115930, 16.4664144722),
    (4.1439360087, -2.9176674704, 0.5246212487),
    (4.0257813209, -2.9025002398, 0.5334261214),
    (3.5118574347, -2.5740236523, 0.5050097282),
    (2.4398158400, -1.7586675341, 0.4191290613),
    (1.9779835097, -1.3337358510, 0.3772169049),
    (1.9559726949, -1.3091355170, 0.3746734515),
    (1.9282822454, -1.2823649693, 0.3704626545),
    (1.9220135179, -1.2812524618, 0.3707011753),
    (1.8942192942, -1.2613293407, 0.3676616051),
)
# src/composite_FP32.cu:86,94 : scale factor 1.001 applied to steps 0..7
FP32_SCALE_FACTOR = 1.001
FP32_SCALED_STEPS = 8

# src/composite_FP16.cu:93-101 ("polar express with refinement")
COEFF_FP16_TABLE = (
    (8.2885332412, -22.5927099246, 15.8201383114),
    (4.1666196466, -2.9679004036, 0.5307623217),
    (4.0611848147, -2.9698947955, 0.5492133813),
    (3.6678301399, -2.7561018955, 0.5421513305),
    (2.7632556383, -2.0607754898, 0.4695405857),
    (2.0527445797, -1.4345145882, 0.4070669182),
    (1.8804816691, -1.2583997294, 0.3779501813),
)
# src/composite_FP16.cu:103,111-113 : scale factor 1.01 applied to all steps
FP16_SCALE_FACTOR = 1.01
FP16_SCALED_STEPS = 7


def step_coefficients(table, scale_factor, n_scaled):
    """Per-step (a, b, c) exactly as the reference computes them in float32.

    src/composite_FP32.cu:86-99 and src/composite_FP16.cu:103-113:
    a /= sf; b /= sf*sf*sf; c /= sf*sf*sf*sf*sf, all in float arithmetic
    (C++ evaluates the products left to right).
    """
    f = np.float32
    sf = f(scale_factor)
    out = []
    for i, (a, b, c) in enumerate(table):
        a, b, c = f(a), f(b), f(c)
        if i < n_scaled:
            a = f(a / sf)
            b = f(b / f(f(sf * sf) * sf))
            c = f(c / f(f(f(f(sf * sf) * sf) * sf) * sf))
        out.append((a, b, c))
    return out


def fp32_chain_coefficients():
    return step_coefficients(COEFF_FP32_TABLE, FP32_SCALE_FACTOR, FP32_SCALED_STEPS)


def fp16_chain_coefficients():
    return step_coefficients(COEFF_FP16_TABLE, FP16_SCALE_FACTOR, FP16_SCALED_STEPS)


# ----------------------------------------------------------------------------
# Scalar transfer function: composite(X) = Q g(L) Q^T in exact arithmetic.
# ----------------------------------------------------------------------------

def polynomial_chain_scalar(x, coeffs):
    """P(x) = p_K o ... o p_1 (x), p_i(x) = a x + b x^3 + c x^5, float64.

    Scalar form of the loop src/composite_FP32.cu:89-119 (A2 = A A,
    A3 = A2 A, A <- a A + b A3 + c A3 A2)."""
    x = np.asarray(x, dtype=np.float64)
    for a, b, c in coeffs:
        x2 = x * x
        x3 = x2 * x
        x = float(a) * x + float(b) * x3 + float(c) * (x3 * x2)
    return x


def transfer_function(x, coeffs):
    """g(x) = x (1 + P(x)) / 2: src/composite_FP32.cu:143-150."""
    x = np.asarray(x, dtype=np.float64)
    return x * (1.0 + polynomial_chain_scalar(x, coeffs)) * 0.5


# ----------------------------------------------------------------------------
# Rounding formats used as GEMM operands (emulated on the CPU).
# ----------------------------------------------------------------------------

def round_f16(x):
    """float32 -> IEEE binary16 round-to-nearest-even -> float32
    (src/utils.cu:106-121 uses __float22half2_rn)."""
    return np.asarray(x, dtype=np.float32).astype(np.float16).astype(np.float32)


def round_bf16(x):
    """float32 -> bfloat16 round-to-nearest-even -> float32 (finite inputs)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


F16X3_SCALE = 256.0  # 2**8: s*x split into hi + lo binary16 keeps lo out of the subnormal range


def round_e4m3(x):
    """float32 -> nearest FP8 E4M3FN value (OCP: 3 mantissa bits, exponent
    bias 7, normal exponents >= -6, subnormal quantum 2**-9, max 448, values
    beyond saturate), ties to even -> float32.  Not a reference format (the
    reference has no FP8 path); the FP8 precision study asks for it."""
    x = np.asarray(x, dtype=np.float32)
    a = np.abs(x).astype(np.float64)
    out = np.empty_like(x)
    big = ~(a < 448.0)
    with np.errstate(divide="ignore"):
        e = np.floor(np.log2(np.where(a > 0, a, 1.0)))
    # np.log2 can land one ulp low at exact powers of two: fix the exponent
    e = np.where(2.0 ** (e + 1) <= a, e + 1, e)
    q = 2.0 ** (np.maximum(e, -6.0) - 3.0)
    r = np.rint(x.astype(np.float64) / q) * q  # np.rint: ties to even
    out[...] = r.astype(np.float32)
    out[big] = np.copysign(np.float32(448.0), x[big])
    out[np.isnan(x)] = np.nan
    return out


MX_BLOCK = 32


def mx_e4m3(x, axis):
    """FP8 e4m3 values with one power-of-two scale per 32 consecutive entries
    along ``axis`` (MX blocks, zero-padded at the end), returned dequantized
    as float32.  A block whose largest magnitude has float32 exponent field E
    gets the E8M0 scale byte b = max(E - 7, 0) (scale 2^(b - 127), so every
    scaled value is below 256: no saturation), an all-zero block b = 127; each
    value is round_e4m3(x / scale) * scale.  No reference counterpart: the
    library's FMT_F16MX cross-term pieces (its own format)."""
    x = np.asarray(x, dtype=np.float32)
    xm = np.moveaxis(x, axis, -1)
    n = xm.shape[-1]
    pad = (-n) % MX_BLOCK
    xp = np.pad(xm, [(0, 0)] * (xm.ndim - 1) + [(0, pad)])
    blk = xp.reshape(xp.shape[:-1] + (-1, MX_BLOCK))
    m = np.max(np.abs(blk), axis=-1, keepdims=True).astype(np.float32)
    E = (m.view(np.uint32) >> 23).astype(np.int64)
    b = np.where(m == 0, 127, np.maximum(E - 7, 0))
    scaled = blk.astype(np.float64) * np.exp2((127 - b).astype(np.float64))
    q = round_e4m3(scaled.astype(np.float32)).astype(np.float64) * np.exp2((b - 127).astype(np.float64))
    out = q.astype(np.float32).reshape(xp.shape)[..., :n]
    return np.moveaxis(out, -1, axis)


def mx_split(x, axis):
    """FMT_F16MX pieces of x: w = 2^8 x, hi = f16(w) (binary16 MFMA operand),
    and the MX e4m3 copies of w and of lo = w - hi (blocks along ``axis``,
    the k index of the product the operand enters)."""
    w = np.asarray(x, dtype=np.float32) * np.float32(F16X3_SCALE)
    hi = round_f16(w)
    return [hi, mx_e4m3(w, axis), mx_e4m3(w - hi, axis)]


def f16x3_value(x):
    """float32 value of a binary16x3 operand entry from its pieces,
    (hi + lo) 2^-8 (within 2^-22 |x| + 2^-33 of x)."""
    hi, lo = split_operand(x, "f16x3")
    return ((hi + lo) * np.float32(1.0 / F16X3_SCALE)).astype(np.float32)


def split_operand(x, fmt):
    """Return the list of float32 pieces a GEMM operand is stored as."""
    x = np.asarray(x, dtype=np.float32)
    if fmt == "f64":
        return [np.asarray(x, dtype=np.float64)]
    if fmt == "f32":
        return [x]
    if fmt == "f16":
        return [round_f16(x)]
    if fmt == "f16x3":
        w = x * np.float32(F16X3_SCALE)
        hi = round_f16(w)
        return [hi, round_f16(w - hi)]
    if fmt == "bf16":
        return [round_bf16(x)]
    if fmt == "e4m3":
        # per-tensor scale 2^8 (operands are bounded by 1 after the spectral bound)
        return [round_e4m3(x * np.float32(F16X3_SCALE))]
    if fmt == "bf16x3":
        hi = round_bf16(x)
        return [hi, round_bf16(x - hi)]
    if fmt in ("bf16x9", "bf16x6"):
        hi = round_bf16(x)
        r = x - hi
        mid = round_bf16(r)
        return [hi, mid, round_bf16(r - mid)]
    raise ValueError(fmt)


def emulated_product(x, y, fmt):
    """x @ y with operands rounded to ``fmt`` pieces and float32 accumulation
    (products of two 16-bit operands are exact in float32)."""
    if fmt == "f64":
        return np.asarray(x, np.float64) @ np.asarray(y, np.float64)
    if fmt == "f16mx":  # MX blocks along k: columns of x, rows of y
        px, py = mx_split(x, 1), mx_split(y, 0)
    else:
        px, py = split_operand(x, fmt), split_operand(y, fmt)
    mm = lambda a, b: (a @ b).astype(np.float32)
    if fmt in ("f32", "f16", "bf16"):
        return mm(px[0], py[0])
    if fmt == "e4m3":
        return mm(px[0], py[0]) * np.float32(1.0 / (F16X3_SCALE * F16X3_SCALE))
    if fmt == "f16x3":
        acc = mm(px[1], py[0]) + mm(px[0], py[1])
        return (acc + mm(px[0], py[0])) * np.float32(1.0 / (F16X3_SCALE * F16X3_SCALE))
    if fmt == "f16mx":
        acc = mm(px[1], py[2]) + mm(px[2], py[1])  # cross terms from the e4m3 copies
        return (acc + mm(px[0], py[0])) * np.float32(1.0 / (F16X3_SCALE * F16X3_SCALE))
    if fmt == "bf16x3":
        return mm(px[0], py[0]) + (mm(px[0], py[1]) + mm(px[1], py[0]))
    if fmt == "bf16x6":
        pairs = [(1, 1), (0, 2), (2, 0), (0, 1), (1, 0), (0, 0)]
    else:  # bf16x9
        pairs = [(2, 2), (1, 2), (2, 1), (1, 1), (0, 2), (2, 0), (0, 1), (1, 0), (0, 0)]
    acc = np.zeros((x.shape[0], y.shape[1]), np.float32)
    for i, j in pairs:
        acc = acc + mm(px[i], py[j])
    return acc


def _sym(m):
    return (m + m.T) * m.dtype.type(0.5)


def _mirror_upper(m):
    """Upper triangle (with the diagonal) copied to the lower triangle."""
    u = np.triu(m)
    return u + np.triu(m, 1).T


def sym_product(x, y, fmt, how="average"):
    """Symmetric result of x @ y (x, y symmetric, commuting in exact
    arithmetic).  how="average": (C + C^T)/2 of the full product;
    how="mirror": the upper triangle of C mirrored (what a half-triangle
    kernel writes); how="full": C itself."""
    c = emulated_product(x, y, fmt)
    if how == "mirror":
        return _mirror_upper(c)
    if how == "average":
        return _sym(c)
    return c


def composite_chain(X, coeffs, step_formats, final_format=None, sym_mode="reference"):
    """Composite polynomial filter chain on an already scaled symmetric X.

    Op order follows src/composite_FP32.cu:89-156 (and the FP16 variant
    src/composite_FP16.cu:106-233; the emulated variant
    src/composite_FP32_emulated.cu:93-227):
        for each step: A2 = A A; A3 = A2 A; A = a A + b A3 + c (A3 A2);
                       A = (A + A^T)/2
        A = (I + A)/2; W = X A; W = (W + W^T)/2
    ``step_formats[i]`` is the operand format of step i's three products;
    "f64" runs the whole chain in float64 (the exact-arithmetic restatement),
    any other format keeps the iterate in float32 like the reference.
    ``sym_mode`` "reference" symmetrizes as the reference does (A and W
    averaged with their transposes, A2/A3 left as computed); "mirror"
    symmetrizes every product by copying its upper triangle (half-triangle
    products); "average" averages every product with its transpose.
    """
    exact = all(f == "f64" for f in step_formats) and final_format in (None, "f64")
    dt = np.float64 if exact else np.float32
    A = np.asarray(X, dtype=dt).copy()
    Xs = A.copy()
    how = {"reference": "full", "mirror": "mirror", "average": "average"}[sym_mode]
    prod = lambda x, y, f: sym_product(x, y, f, how).astype(dt)
    for (a, b, c), fmt in zip(coeffs, step_formats):
        a, b, c = dt(a), dt(b), dt(c)
        A2 = prod(A, A, fmt)
        A3 = prod(A2, A, fmt)
        # a binary16x3 step reads A^3 for the b term back from its two
        # operand pieces, (hi + lo) 2^-8, instead of a float32 copy
        A3b = f16x3_value(A3) if fmt == "f16x3" else A3
        A = a * A + b * A3b
        A = A + c * prod(A3, A2, fmt)
        if sym_mode == "reference":
            A = _sym(A)
    n = A.shape[0]
    A = (A + np.eye(n, dtype=dt)) * dt(0.5)
    W = prod(Xs, A, final_format or step_formats[-1])
    if sym_mode == "reference":
        W = _sym(W)
    return W.astype(np.float64)


def composite_FP32(X):
    """src/composite_FP32.cu:15-164 (FP32 chain, 10 steps)."""
    c = fp32_chain_coefficients()
    return composite_chain(X, c, ["f32"] * len(c), "f32")


def composite_FP16(X):
    """src/composite_FP16.cu:16-246 (FP16 operands, FP32 accumulate, 7 steps)."""
    c = fp16_chain_coefficients()
    return composite_chain(X, c, ["f16"] * len(c), "f16")


def composite_FP32_emulated(X):
    """src/composite_FP32_emulated.cu:20-235 restated with the published BF16x9
    split product (Henry, Tang, Heinecke, ARITH 2019)."""
    c = fp32_chain_coefficients()
    return composite_chain(X, c, ["bf16x9"] * len(c), "bf16x9")


def composite_exact(X, table="fp32"):
    """The chain in float64 arithmetic (same op order, coefficients rounded to
    float like the reference)."""
    c = fp32_chain_coefficients() if table == "fp32" else fp16_chain_coefficients()
    return composite_chain(X, c, ["f64"] * len(c), "f64")


# ----------------------------------------------------------------------------
# Lanczos spectral-norm bound: src/lanczos.cu:41-293
# ----------------------------------------------------------------------------

def lanczos_two_norm(A, max_iter=20, tol=1e-10, q0=None, return_state=False, keep_exact=True, mode="reference"):
    """Restatement of approximate_two_norm (src/lanczos.cu:41-293).

    Lanczos on A^T A without reorthogonalisation.  Break rule of
    src/lanczos.cu:145 (beta_k <= tol*alpha_k and k > 1, alpha_k dropped) and
    :152-159 (beta_k == 0).  T = tridiag(alpha, beta) (src/lanczos.cu:16-38),
    theta = largest Ritz value (:231), y = V u (:242).  The residual follows
    the reference as written: ``cublasDcopy(n, ry, 1, w1, 1)`` at :255 copies
    ry INTO w1, so r = A y - theta y (not A^T A y - theta y).
    lo = sqrt(theta) (:264), up = sqrt(theta + ||r||) (:270).  The all-zero
    path (nb_iter == 0) returns lo = 0, up = 1 (:171-184).
    ``q0`` replaces the cuRAND start vector (src/utils.cu:201-214); it is
    normalised as at :103-107.
    mode="tight" is the library's opt-in bound (no reference counterpart): the
    residual of the Ritz pair of A^T A itself, r = A (A y) - theta y, so that
    up - ||A||_2 shrinks with the Ritz error.
    """
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    max_iter = min(int(max_iter), n)
    if q0 is None:
        q0 = np.random.default_rng(0).random(n) + 1e-300
    q = np.array(q0, dtype=np.float64)
    nq = np.linalg.norm(q)
    if nq != 0.0:
        q = q * (1.0 / nq)
    V = np.zeros((n, max_iter))
    V[:, 0] = q
    v_old = np.zeros(n)
    alpha = np.zeros(max_iter)
    beta = np.zeros(max_iter)
    beta_old = 0.0
    nb = 0
    for k in range(max_iter):
        w = A.T @ (A @ q)
        alpha[k] = q @ w
        w = w - alpha[k] * q - beta_old * v_old
        beta[k] = np.linalg.norm(w)
        if beta[k] <= tol * alpha[k] and k > 1:
            # with keep_exact, iteration n - 1 of an n x n matrix is kept: the
            # Krylov space is the whole space and beta is rounding noise
            # (the reference drops it, which can leave up < ||A|| for n <= 20)
            if keep_exact and k == n - 1 and alpha[k] > 0.0:
                nb += 1
            break
        v_old = q.copy()
        if beta[k] == 0.0:
            # exact invariant subspace: the reference drops iteration k here
            # (src/lanczos.cu:152-159), which turns e.g. any 1x1 matrix into
            # the "zero matrix" case; with keep_exact it is kept (its Ritz
            # value is exact) unless alpha_k == 0.
            if keep_exact and alpha[k] > 0.0:
                nb += 1
            break
        q = w * (1.0 / beta[k])
        if k < max_iter - 1:
            V[:, k + 1] = q
        beta_old = beta[k]
        nb += 1
    if nb == 0:
        res = (0.0, 1.0)
        if return_state:
            return res + (dict(nb=0, alpha=alpha, beta=beta, theta=None),)
        return res
    T = np.diag(alpha[:nb]) + np.diag(beta[: nb - 1], 1) + np.diag(beta[: nb - 1], -1)
    evals, evecs = np.linalg.eigh(T)
    theta = evals[-1]
    u = evecs[:, -1]
    y = V[:, :nb] @ u
    if mode == "tight":
        r = A.T @ (A @ y) - theta * y
    elif mode == "reference":
        r = A @ y - theta * y
    else:
        raise ValueError(mode)
    lo = np.sqrt(theta)
    up = np.sqrt(theta + np.linalg.norm(r))
    if return_state:
        return lo, up, dict(nb=nb, alpha=alpha[:nb], beta=beta[:nb], theta=theta, T=T)
    return lo, up


def composite_auto_scale(X, method="fp32", q0=None, max_iter=20, tol=1e-10):
    """composite_*_auto_scale (src/composite_FP32.cu:166-191,
    src/composite_FP16.cu:248-274, src/composite_FP32_emulated.cu:237-262):
    up from Lanczos, s = up if up > 0 else 1, Y = s * chain(X / s)."""
    _, up = lanczos_two_norm(X, max_iter, tol, q0)
    s = up if up > 0.0 else 1.0
    Xs = np.asarray(X, np.float64) * (1.0 / s)
    f = {"fp32": composite_FP32, "fp16": composite_FP16,
         "fp32_emulated": composite_FP32_emulated,
         "exact": composite_exact}[method]
    return f(Xs) * s


# ----------------------------------------------------------------------------
# EVD projection (oracle and CPU-baseline method): src/eig_FP64_psd.cu:24-85
# ----------------------------------------------------------------------------

def eig_FP64_psd(X, return_eigenvalues=False):
    """dsyevd (jobz=V, uplo=U) -> clamp at 0 -> scale columns -> V diag V^T.
    LAPACK dsyevd via scipy (driver 'evd', upper triangle)."""
    import scipy.linalg as sla
    X = np.asarray(X, dtype=np.float64)
    w, V = sla.eigh(X, lower=False, driver="evd", check_finite=False)
    wp = np.maximum(w, 0.0)
    Y = (V * wp) @ V.T
    if return_eigenvalues:
        return Y, w
    return Y


def _scipy_lapack():
    """ctypes handle on the LAPACK that scipy.linalg uses (its bundled
    OpenBLAS, whose symbols carry the prefix scipy_)."""
    import ctypes
    import scipy.linalg  # noqa: F401
    from threadpoolctl import threadpool_info
    paths = [i["filepath"] for i in threadpool_info()
             if i.get("internal_api") == "openblas" and "scipy.libs" in i.get("filepath", "")]
    if not paths:
        raise RuntimeError("scipy's OpenBLAS not found")
    return ctypes.CDLL(paths[0]), paths[0]


def eig_FP64_psd_staged(A_in, log=None):
    """eig_FP64_psd (src/eig_FP64_psd.cu:24-85) on the host: LAPACK dsyevd
    (jobz = V, uplo = U) run as its three stages -- dsytrd (reduction to
    tridiagonal form), dstedc (divide and conquer with eigenvectors), dormtr
    (back-transformation) -- with the optimal workspaces LAPACK reports, from
    the LAPACK that scipy.linalg uses; then the clamp and V diag(max(w, 0)) V^T
    (src/eig_FP64_psd.cu:51-75).  This is the timed CPU baseline; cuSOLVER's
    Dsyevd also runs with its queried workspace (src/eig_FP64_psd.cu:34-43).
    log(msg) receives each stage's time when it ends.  Returns (Y, w ascending,
    stage seconds)."""
    import ctypes
    import numpy as np
    import time
    L, _ = _scipy_lapack()
    n = A_in.shape[0]
    A = np.array(A_in, dtype=np.float64, order="F", copy=True)
    c_int, c_dbl = ctypes.c_int, ctypes.c_double
    P_ = ctypes.POINTER
    by = ctypes.byref
    N = c_int(n)
    info = c_int(0)
    U, I, Lc, Nc = (ctypes.c_char(c) for c in (b"U", b"I", b"L", b"N"))
    one = ctypes.c_size_t(1)
    dptr = lambda a: a.ctypes.data_as(P_(c_dbl))
    d = np.empty(n)
    e = np.empty(max(n - 1, 1))
    tau = np.empty(max(n - 1, 1))
    q = np.empty(1)
    stages = {}
    say = log if log is not None else (lambda msg: None)

    def check(name):
        if info.value != 0:
            raise RuntimeError(f"LAPACK {name} info {info.value}")

    L.scipy_dsytrd_(by(U), by(N), dptr(A), by(N), dptr(d), dptr(e), dptr(tau), dptr(q), by(c_int(-1)), by(info), one)
    work = np.empty(max(int(q[0]), 1))
    t = time.perf_counter()
    L.scipy_dsytrd_(by(U), by(N), dptr(A), by(N), dptr(d), dptr(e), dptr(tau), dptr(work), by(c_int(work.size)),
                    by(info), one)
    stages["dsytrd"] = time.perf_counter() - t
    check("dsytrd")
    say(f"eig_FP64_psd_staged n={n}: dsytrd {stages['dsytrd']:.1f} s")
    Z = np.empty((n, n), order="F")
    qi = c_int(0)
    L.scipy_dstedc_(by(I), by(N), dptr(d), dptr(e), dptr(Z), by(N), dptr(q), by(c_int(-1)), by(qi), by(c_int(-1)),
                    by(info), one)
    work = np.empty(max(int(q[0]), 1))
    iwork = np.empty(max(qi.value, 1), dtype=np.int32)
    t = time.perf_counter()
    L.scipy_dstedc_(by(I), by(N), dptr(d), dptr(e), dptr(Z), by(N), dptr(work), by(c_int(work.size)),
                    iwork.ctypes.data_as(P_(c_int)), by(c_int(iwork.size)), by(info), one)
    stages["dstedc"] = time.perf_counter() - t
    check("dstedc")
    say(f"eig_FP64_psd_staged n={n}: dstedc {stages['dstedc']:.1f} s")
    L.scipy_dormtr_(by(Lc), by(U), by(Nc), by(N), by(N), dptr(A), by(N), dptr(tau), dptr(Z), by(N), dptr(q),
                    by(c_int(-1)), by(info), one, one, one)
    work = np.empty(max(int(q[0]), 1))
    t = time.perf_counter()
    L.scipy_dormtr_(by(Lc), by(U), by(Nc), by(N), by(N), dptr(A), by(N), dptr(tau), dptr(Z), by(N), dptr(work),
                    by(c_int(work.size)), by(info), one, one, one)
    stages["dormtr"] = time.perf_counter() - t
    check("dormtr")
    say(f"eig_FP64_psd_staged n={n}: dormtr {stages['dormtr']:.1f} s")
    del A, work
    t = time.perf_counter()
    Y = (Z * np.maximum(d, 0.0)) @ Z.T
    stages["clamp_gemm"] = time.perf_counter() - t
    say(f"eig_FP64_psd_staged n={n}: clamp + V diag V^T {stages['clamp_gemm']:.1f} s")
    order = np.argsort(d)  # dstedc returns ascending eigenvalues already
    return Y, d[order], stages


def eig_FP32_psd(X, return_eigenvalues=False):
    """src/eig_FP32_psd.cu:23-96: FP64 -> FP32, ssyevd, clamp, FP32 V diag V^T."""
    import scipy.linalg as sla
    Xs = np.asarray(X, dtype=np.float32)
    w, V = sla.eigh(Xs, lower=False, driver="evd", check_finite=False)
    w = w.astype(np.float32)
    Y = ((V * np.maximum(w, np.float32(0))) @ V.T).astype(np.float64)
    if return_eigenvalues:
        return Y, w.astype(np.float64)
    return Y


def closed_form_projection(Q, lam):
    """Pi(X) = Q max(L,0) Q^T for X = Q diag(lam) Q^T."""
    Q = np.asarray(Q, np.float64)
    return (Q * np.maximum(np.asarray(lam, np.float64), 0.0)) @ Q.T


def rel_fro_error(Y, P):
    Y = np.asarray(Y, np.float64)
    P = np.asarray(P, np.float64)
    return float(np.linalg.norm(Y - P) / np.linalg.norm(P))


# ----------------------------------------------------------------------------
# Extremal eigenpairs and deflated projections
# ----------------------------------------------------------------------------

def lanczos_eigenpairs(A, k, q0, max_iter=0, ortho_tol=1e-5, upper_bound_only=False, residual_row="first",
                       normalize=False):
    """Restatement of compute_eigenpairs (src/lanczos.cu:319-547).

    Lanczos on A with the whole basis kept, no reorthogonalisation
    (:359-389): w = A q_m; w -= beta_{m-1} q_{m-1}; alpha_m = q_m . w;
    w -= alpha_m q_m; beta_m = ||w||; stop if beta_m <= DBL_EPSILON (the
    iteration counts); q_{m+1} = w * (1/beta_m).  max_iter = 0 means n
    (:331-332); max_iter may exceed n.  T = tridiag(alpha, beta[:nb-1]),
    eigenpairs ascending (:391-423).  Residuals |beta_last Z(row, j)|, row 0
    as the reference reads Z[j*m] (:295-301), or the last row.  Order by
    decreasing |theta| (:438-447); keep a pair when res < ortho_tol and
    |x_a . x| <= ortho_tol for every kept x_a, x = Q z unnormalised, until k
    (:453-507).  bound = |theta_max| + res_max + 1e-3 (:511-547).
    normalize divides the kept vectors by their norms after the selection
    (not in the reference; the deflated projection needs unit vectors).
    Returns (values, vectors (n x r), bound, state).
    """
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    max_iter = n if int(max_iter) == 0 else int(max_iter)
    q = np.array(q0, dtype=np.float64)
    nq = np.linalg.norm(q)
    if nq != 0.0:
        q = q * (1.0 / nq)
    Q = np.zeros((n, max_iter + 1))
    Q[:, 0] = q
    alpha, beta = [], []
    for m in range(max_iter):
        w = A @ Q[:, m]
        if m > 0:
            w = w - beta[-1] * Q[:, m - 1]
        a = Q[:, m] @ w
        alpha.append(a)
        w = w - a * Q[:, m]
        b = np.linalg.norm(w)
        beta.append(b)
        if b <= np.finfo(np.float64).eps:
            break
        Q[:, m + 1] = w * (1.0 / b)
    nb = len(alpha)
    T = np.diag(alpha) + np.diag(beta[: nb - 1], 1) + np.diag(beta[: nb - 1], -1)
    theta, Z = np.linalg.eigh(T)
    row = 0 if residual_row == "first" else nb - 1
    res = np.abs(beta[-1] * Z[row, :])
    idx = sorted(range(nb), key=lambda j: -abs(theta[j]))
    bound = abs(theta[idx[0]]) + res[idx[0]] + 1e-3
    state = dict(nb=nb, alpha=np.array(alpha), beta=np.array(beta), theta=theta, Z=Z, res=res, Q=Q[:, :nb])
    if upper_bound_only:
        return None, None, bound, state
    vals, vecs = [], []
    for j in idx:
        if len(vals) >= k:
            break
        if not res[j] < ortho_tol:
            continue
        x = Q[:, :nb] @ Z[:, j]
        if vecs and np.max(np.abs(np.array(vecs) @ x)) > ortho_tol:
            continue
        vals.append(theta[j])
        vecs.append(x)
    V = np.array(vecs).T if vecs else np.zeros((n, 0))
    if normalize and vecs:
        V = V / np.linalg.norm(V, axis=0)
    return np.array(vals), V, bound, state


def lobpcg(A, m, X0, maxiter=100, tol=1e-8, warm=False, sign=1.0):
    """Restatement of lobpcg (src/lobpcg.cu:12-256) for the m largest
    eigenpairs of sign*A.

    Start: X0 orthonormalised by Householder QR unless warm (:84-104), then
    Rayleigh-Ritz on it (:106-133).  Iteration (:135-218): R = AX - X lam
    (X lam rounded first, as Ddgmm + Daxpy), stop when ||R||_F < tol;
    S = [X, R, dX]; Q = Householder QR of S; T = Q^T A Q symmetrised;
    the m largest Ritz pairs, descending: X' = Q Z_m, dX = X' - X.
    Stated differences, shared with the library and not changing the limit:
    the first iteration stacks [X, R] (the reference's [X, R, X] adds a
    dependent block), the initial Rayleigh-Ritz pairs vectors with their own
    values, and AX' is carried as (A Q) Z_m.
    Returns (V (n x m), lam (m, descending), iterations, ||R||_F).
    """
    A = sign * np.asarray(A, dtype=np.float64)
    X = np.array(X0, dtype=np.float64)
    if not warm:
        X, _ = np.linalg.qr(X, mode="reduced")

    def rayleigh_ritz(Qb):
        AQ = A @ Qb
        T = Qb.T @ AQ
        T = 0.5 * T + 0.5 * T.T
        th, Z = np.linalg.eigh(T)
        Zm = Z[:, ::-1][:, :m]
        return Qb @ Zm, AQ @ Zm, th[::-1][:m].copy()

    X, AX, lam = rayleigh_ritz(X)
    dX = None
    iters, rn = 0, 0.0
    for it in range(1, int(maxiter) + 1):
        R = AX - X * lam
        rn = float(np.linalg.norm(R))
        if rn < tol:
            break
        iters = it
        S = np.hstack([X, R] if dX is None else [X, R, dX])
        Qs, _ = np.linalg.qr(S, mode="reduced")
        Xn, AXn, lam = rayleigh_ritz(Qs)
        dX = Xn - X
        X, AX = Xn, AXn
    return X, lam, iters, rn


def _rank_update(X, V, c):
    """X + V diag(c) V^T (the Dger loops of the deflated projections)."""
    if V.shape[1] == 0:
        return X
    return X + (V * c) @ V.T


def composite_FP32_auto_scale_deflate(X, k=40, ortho_tol=1e-3, q0=None, q0_bound=None, residual_row="first",
                                      chain=None, normalize=True):
    """src/composite_FP32.cu:193-254: the k largest-|lambda| Lanczos pairs
    (compute_eigenpairs, max_iter = n) removed; bound of the remainder by
    compute_eigenpairs(upper_bound_only, max_iter 100, ortho 1e-5); s =
    1.1 up + 1e-5 if up > 0 else 1; s * composite_FP32(X / s); positive
    pairs added back.  k = 0 deflates nothing.  normalize=False keeps the
    reference's unnormalised Ritz vectors (src/lanczos.cu:478-481), whose
    x x^T removes theta ||x||^2 once the basis has lost orthogonality."""
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[0]
    q0 = np.ones(n) if q0 is None else q0
    q0_bound = q0 if q0_bound is None else q0_bound
    if k > 0:
        vals, V, _, _ = lanczos_eigenpairs(X, k, q0, 0, ortho_tol, residual_row=residual_row, normalize=normalize)
    else:
        vals, V = np.zeros(0), np.zeros((n, 0))
    Xd = _rank_update(X, V, -vals)
    _, _, up, _ = lanczos_eigenpairs(Xd, 0, q0_bound, 100, 1e-5, upper_bound_only=True, residual_row=residual_row)
    s = 1.1 * up + 1e-5 if up > 0.0 else 1.0
    f = composite_FP32 if chain is None else chain
    Y = f(Xd * (1.0 / s)) * s
    return _rank_update(Y, V, np.maximum(vals, 0.0))


def _lobpcg_deflate(X, k, maxiter, tol, X0):
    n = X.shape[0]
    Vmax, lmax, _, _ = lobpcg(X, k, X0, maxiter, tol)
    X = _rank_update(X, Vmax, -lmax)
    Vmin, mu, _, _ = lobpcg(X, k, X0, maxiter, tol, sign=-1.0)
    X = _rank_update(X, Vmin, mu)
    return X, np.hstack([Vmax, Vmin]), np.concatenate([lmax, -mu])


def composite_FP16_auto_scale_deflate(X, k=40, maxiter=100, tol=1e-10, X0=None, q0=None, chain=None):
    """src/composite_FP16.cu:276-380: LOBPCG removes the k largest pairs,
    then the k smallest of the remainder (largest of its negation, :316-333);
    approximate_two_norm bound, s = 1.1 up + 1e-5 if up > 0 else 1;
    s * composite_FP16(X / s); positive pairs of both sets added back."""
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[0]
    X0 = np.random.default_rng(0).random((n, k)) if X0 is None else X0
    Xd, V, lam = _lobpcg_deflate(X, k, maxiter, tol, X0)
    _, up = lanczos_two_norm(Xd, 20, 1e-10, q0)
    s = 1.1 * up + 1e-5 if up > 0.0 else 1.0
    f = composite_FP16 if chain is None else chain
    Y = f(Xd * (1.0 / s)) * s
    return _rank_update(Y, V, np.maximum(lam, 0.0))


def eig_FP64_deflate(X, k=40, maxiter=100, tol=1e-10, X0=None):
    """src/eig_FP64_psd.cu:87-177: the LOBPCG deflation of
    composite_FP16_auto_scale_deflate, eig_FP64_psd of the remainder,
    positive pairs added back."""
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[0]
    X0 = np.random.default_rng(0).random((n, k)) if X0 is None else X0
    Xd, V, lam = _lobpcg_deflate(X, k, maxiter, tol, X0)
    return _rank_update(eig_FP64_psd(Xd), V, np.maximum(lam, 0.0))


# ----------------------------------------------------------------------------
# Input generators
# ----------------------------------------------------------------------------

class GlibcRandom:
    """glibc random()/rand() TYPE_3 additive-feedback generator (the published
    glibc random_r algorithm: r[i] = r[i-3] + r[i-31] mod 2^32, output r >> 1,
    seeded by r[i] = 16807 r[i-1] mod (2^31-1), 310 warm-up outputs)."""

    RAND_MAX = 2147483647

    def __init__(self, seed=1):
        r = [0] * 34
        r[0] = seed if seed != 0 else 1
        for i in range(1, 31):
            hi, lo = divmod(r[i - 1], 127773)
            word = 16807 * lo - 2836 * hi
            if word < 0:
                word += 2147483647
            r[i] = word
        for i in range(31, 34):
            r[i] = r[i - 31]
        self._r = r
        self._warm = 310

    def take(self, count):
        """Return ``count`` successive rand() outputs as int64 array."""
        total = self._warm + count
        r = np.zeros(34 + total, dtype=np.uint64)
        r[:34] = np.array(self._r, dtype=np.uint64)
        mask = np.uint64(0xFFFFFFFF)
        # r[i] = r[i-31] + r[i-3]; vectorised in chunks of 3 (the shortest lag)
        i = 34
        end = 34 + total
        while i < end:
            j = min(i + 3, end)
            r[i:j] = (r[i - 31:j - 31] + r[i - 3:j - 3]) & mask
            i = j
        out = (r[34 + self._warm:] >> np.uint64(1)).astype(np.int64)
        self._r = [int(v) for v in r[-34:]]
        self._warm = 0
        return out


def generate_and_project(n, radius=1.0):
    """test/test_utils.hpp:20-164 (generateAndProject) on the host.

    srand(1); M_ij = rand()/RAND_MAX - 0.5 (filled in memory order, i.e.
    column-major); Q from Householder QR (LAPACK dgeqrf/dorgqr, the same
    convention cuSOLVER follows); lambda_i = -r + 2 i r/(n-1); A = (Q L) Q^T,
    symmetrised; oracle = FP64 EVD projection.  Returns (A, A_psd, Q, lam)."""
    g = GlibcRandom(1)
    vals = g.take(n * n).astype(np.float64) / GlibcRandom.RAND_MAX - 0.5
    M = vals.reshape(n, n, order="F")
    Q, _ = np.linalg.qr(M, mode="reduced")
    lam = -radius + 2.0 * np.arange(n) * radius / float(n - 1)
    A = (Q * lam) @ Q.T
    A = 0.5 * (A + A.T)
    return A, eig_FP64_psd(A), Q, lam


def haar_orthogonal(n, seed):
    """Haar-distributed Q: QR of a seeded N(0,1) matrix with the sign fix
    Q <- Q diag(sign(diag R))."""
    G = np.random.default_rng(seed).standard_normal((n, n))
    Q, R = np.linalg.qr(G)
    return Q * np.sign(np.diag(R))


def uniform512_input(n=512):
    """The n=512 case: Q from QR of N(0,1) seed 1, lambda ~ U[-1,1] seed 1001."""
    Q = haar_orthogonal(n, 1)
    lam = np.random.default_rng(1001).uniform(-1.0, 1.0, n)
    X = (Q * lam) @ Q.T
    X = 0.5 * (X + X.T)
    return X, Q, lam


def clustered_spectrum(n, seed, kind):
    """The clustered and plain spectra of the workloads (scaled by the same rule at any n)."""
    rng = np.random.default_rng(seed + 100)
    if kind == "normal_cluster":
        lam = rng.standard_normal(n)
        lam *= 10.0 / np.max(np.abs(lam))
        idx = rng.choice(n, int(round(0.3 * n)), replace=False)
        lam[idx] = rng.uniform(-1e-2, 1e-2, idx.size)
    elif kind == "uniform_cluster":
        lam = rng.uniform(-5.0, 5.0, n)
        idx = rng.choice(n, int(round(0.1 * n)), replace=False)
        lam[idx] = rng.uniform(-1e-3, 1e-3, idx.size)
    elif kind == "uniform":
        lam = rng.uniform(-1.0, 1.0, n)
    elif kind == "normal":
        lam = rng.standard_normal(n)
    else:
        raise ValueError(kind)
    return lam


def spectral_input(n, seed, kind):
    Q = haar_orthogonal(n, seed)
    lam = clustered_spectrum(n, seed, kind)
    X = (Q * lam) @ Q.T
    X = 0.5 * (X + X.T)
    return X, Q, lam


# How each oracle function is pinned (checked in tests/test_oracle.py).
PINNING = {
    "f16x3_value": "within 2^-22 |x| + 2^-33 of x on [-1, 1] (the split's own bound); exact on values hi + lo "
                   "already on the binary16x3 grid (tests/test_oracle.py)",
    "step_coefficients": "float32 re-derivation + transfer-function KATs",
    "transfer_function": "max|g-ReLU| on [-1,1] (4.48e-6 FP32 table, 4.65e-5 FP16 table), "
                         "oddness of P, g(x)=x for x in [1e-2,1], g(x)=0 for x in [-1,-1e-2]",
    "composite_chain(f64)": "equals Q g(L) Q^T to 1e-12 (closed form); reference KATs "
                            "(3x3 identity map, diag(1,-2,3,-4,5))",
    "emulated_product": "f32 matches float64 within 1e-6; f16/bf16 rounding matches numpy "
                        "float16 / bit patterns; x3/x9 splits reconstruct the operand",
    "lanczos_two_norm": "5x5 KAT (7.089012550562305) bracketed; up formula vs direct "
                        "recomputation from T; lo <= ||A|| on random inputs; zero matrix -> (0,1); "
                        "tight mode: a complete Krylov space gives up = ||A||_2 to 1e-12 (the reference "
                        "residual of the same pair does not), up >= ||A||_2 >= lo on random spectra; "
                        "keep_exact: a 3x3 input whose reference-rule bound undercuts ||A||_2 gets "
                        "lo = ||A||_2 and up >= ||A||_2, larger inputs are unchanged",
    "eig_FP64_psd": "diag KAT, PSD fixed point, closed form Q max(L,0) Q^T",
    "eig_FP64_psd_staged": "equals eig_FP64_psd (scipy eigh, driver evd) and the closed form on random and "
                           "diagonal inputs; eigenvalues equal numpy eigvalsh",
    "GlibcRandom": "first outputs after srand(1) (1804289383, 846930886, ...) and libc rand()",
    "mx_e4m3": "KATs (exact grid values, ties, zero and tiny blocks), brute force of every value against the "
               "nearest e4m3 code times the frexp-derived block scale, axis/transpose and ragged-end padding; "
               "the GPU's v_cvt_pk_fp8_f32 rounding matches round_e4m3 bit for bit",
    "emulated_product(f16mx)": "error band between FP16x3 and binary16 with planted mistakes (no cross terms, "
                               "hi8.hi8 pairing, wrong block axis) outside it; GPU products match to 1e-6",
    "round_e4m3": "OCP E4M3FN grid KATs (448, 2^-6, 2^-9, ties), brute force over all 256 codes, "
                  "torch float8_e4m3fn casts in range",
    "lanczos_eigenpairs": "5x5 KAT (bound >= 7.0890, lanczos_test.hpp:44-84); full-length runs on small "
                          "matrices return the largest-|lambda| eigenpairs of numpy eigh; last-row residual "
                          "equals ||A x - theta x|| and first-row residual equals |beta_last q0 . x| "
                          "(identities of an orthonormal basis); diag(1,-2,3,-4,5) picks 5 then -4",
    "lobpcg": "6x6 KAT top-2 {7.3191, 5.6639} (lobpcg_test.hpp:10-55); top-m of numpy eigh to 1e-8 on "
              "rand()/RAND_MAX symmetric matrices (lobpcg_test.hpp:57-121); sign=-1 gives the smallest",
    "composite_FP32_auto_scale_deflate": "diag(1,-2,3,-4,5), k=2 -> diag(1,0,3,0,5) within 1e-3 "
                                         "(composite_FP32_test.hpp:249-290, k as intended); error vs EVD "
                                         "projection on generateAndProject inputs",
    "composite_FP16_auto_scale_deflate": "rel-F error <= 1e-3 on symmetrised rand() matrices and <= 5e-3 "
                                         "on generateAndProject(radius 10) (composite_FP16_test.hpp)",
    "eig_FP64_deflate": "equals eig_FP64_psd of the input to rounding (deflation is exact algebra)",
}
