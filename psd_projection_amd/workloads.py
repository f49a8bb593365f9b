"""Synthetic benchmark workloads, generated on the GPU.

Every matrix is X = Q diag(lam) Q^T with Q Haar-distributed (QR of a seeded
Gaussian, sign-fixed so that diag(R) > 0), symmetrised; the exact projection
Pi(X) = Q max(lam, 0) Q^T comes with it.  torch is used as plumbing (FP64
QR/GEMM on the device); the projection itself runs in libpsdp.
"""
from __future__ import annotations

import numpy as np


def spectrum(n: int, seed: int, kind: str) -> np.ndarray:
    """Eigenvalues of the benchmark workloads:
    uniform: U[-1,1] (the n=512 case and the 256-block batch);
    normal_cluster: N(0,1) scaled to max|lam| = 10 with 30% replaced by
    U[-1e-2,1e-2]; uniform_cluster: U[-5,5] with 10% replaced by
    U[-1e-3,1e-3]; normal: N(0,1) (the served request stream); even10: evenly spaced in
    [-10, 10], the spectrum of the reference's test generator
    (test/test_utils.hpp:20-164, radius 10)."""
    if kind == "even10":
        return -10.0 + 20.0 * np.arange(n) / float(n - 1)
    rng = np.random.default_rng(seed + 100)
    if kind == "uniform":
        return rng.uniform(-1.0, 1.0, n)
    if kind == "normal":
        return rng.standard_normal(n)
    if kind == "normal_cluster":
        lam = rng.standard_normal(n)
        lam *= 10.0 / np.max(np.abs(lam))
        idx = rng.choice(n, int(round(0.3 * n)), replace=False)
        lam[idx] = rng.uniform(-1e-2, 1e-2, idx.size)
        return lam
    if kind == "uniform_cluster":
        lam = rng.uniform(-5.0, 5.0, n)
        idx = rng.choice(n, int(round(0.1 * n)), replace=False)
        lam[idx] = rng.uniform(-1e-3, 1e-3, idx.size)
        return lam
    raise ValueError(kind)


def haar(n: int, seed: int, device="cuda"):
    import torch
    g = torch.Generator(device=device).manual_seed(int(seed))
    G = torch.randn(n, n, dtype=torch.float64, device=device, generator=g)
    Q, R = torch.linalg.qr(G)
    del G
    Q.mul_(torch.sign(torch.diagonal(R)).unsqueeze(0))
    del R
    return Q


def householder(n: int, seed: int, device="cuda"):
    """Q = I - 2 v v^T / (v^T v) for a seeded Gaussian v: exactly orthogonal and
    built without a QR factorisation (a cheap stand-in for Haar Q where only
    the shape and spectrum of the workload matter, e.g. for counter runs)."""
    import torch
    g = torch.Generator(device=device).manual_seed(int(seed))
    v = torch.randn(n, 1, dtype=torch.float64, device=device, generator=g)
    Q = torch.eye(n, dtype=torch.float64, device=device)
    Q.sub_((2.0 / float((v * v).sum())) * (v @ v.T))
    return Q


def spectral_matrix(n: int, seed: int, kind: str, device="cuda", want_projection=True, basis="haar"):
    """Returns (X float64, Pi(X) float64 or None, lam numpy).  basis: "haar"
    (QR of a seeded Gaussian with the sign fix) or "householder"."""
    import torch
    lam = spectrum(n, seed, kind)
    Q = haar(n, seed, device) if basis == "haar" else householder(n, seed, device)
    L = torch.from_numpy(lam).to(device)
    X = (Q * L) @ Q.T
    X = (X + X.T) * 0.5
    Pm = None
    if want_projection:
        Pm = (Q * L.clamp(min=0)) @ Q.T
    del Q
    return X, Pm, lam


def rel_fro_error(Y, Pm) -> float:
    import torch
    d = (Y.to(torch.float64) - Pm).norm()
    return float(d / Pm.norm())
