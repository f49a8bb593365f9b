"""Writes the golden fixtures under tests/golden/.

reference_kats.json holds the known-answer data of the reference's own tests
(inputs and expected outputs, copied as data) plus values fixed by the
mathematics (numpy/LAPACK eigen-decompositions):
  * composite_FP32_test.hpp:27-51   3x3 PSD input /15 -> itself (tol 1e-3)
  * composite_FP32_test.hpp:258-276 diag(1,-2,3,-4,5) -> diag(1,0,3,0,5) (tol 1e-3)
  * lanczos_test.hpp:20-31          5x5 matrix, ||A||_2 = 7.0890 (lo <= . <= up)
  * lobpcg_test.hpp:23-42           6x6 matrix, top-2 eigenvalues 7.3191, 5.6639
  * lanczos_test.hpp:44-84          5x5 matrix, compute_eigenpairs(k=3) bound >= 7.0890
  * composite_FP32_test.hpp:249-290 the diag(1,-2,3,-4,5) case through the deflated
                                    projection with k = 2 (the test binds 2 to tol and
                                    0 to k by position; k = 2 is what it means)
transfer_*.npz holds the scalar transfer function g(x) of both coefficient
tables on a grid (computed by the oracle; regression data for the GPU tests).
chain_small.npz holds small FP64 chain inputs/outputs (regression data).

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
from oracle import psdp_oracle as O  # noqa: E402


def main():
    kats = {}
    A3 = (np.array([1, 2, 3, 2, 5, 6, 3, 6, 9], dtype=np.float64) / 15.0).tolist()
    kats["composite_fp32_deterministic"] = {
        "source": "test/composite_FP32_test.hpp:18-78",
        "n": 3, "input_colmajor": A3,
        "expected_colmajor": [0.0667, 0.1333, 0.2000, 0.1333, 0.3333, 0.4000, 0.2000, 0.4000, 0.6000],
        "tol": 1e-3, "entry": "composite_FP32 (no scaling)",
    }
    d5 = [1.0, -2.0, 3.0, -4.0, 5.0]
    kats["simple5"] = {
        "source": "test/composite_FP32_test.hpp:249-290",
        "n": 5, "diagonal": d5, "expected_diagonal": [max(x, 0.0) for x in d5], "tol": 1e-3,
        "entry": "auto-scaled projection",
    }
    L5 = [4.0, 1.0, 2.0, 0.0, 0.0, 1.0, 3.0, 0.0, 1.0, 0.0, 2.0, 0.0, 5.0, 1.0, 1.0,
          0.0, 1.0, 1.0, 2.0, 1.0, 0.0, 0.0, 1.0, 1.0, 3.0]
    M5 = np.array(L5).reshape(5, 5, order="F")
    kats["lanczos_5x5"] = {
        "source": "test/lanczos_test.hpp:8-42",
        "n": 5, "input_colmajor": L5, "two_norm_reference": 7.0890,
        "two_norm": float(np.linalg.norm(M5, 2)),
        "eigenvalues": np.linalg.eigvalsh(M5).tolist(),
        "max_iter": 50, "tol": 1e-6,
    }
    L6 = [4.0, 1.0, 2.0, 0.0, 0.0, 0.0, 1.0, 3.0, 0.0, 1.0, 0.0, 0.0, 2.0, 0.0, 5.0, 1.0, 1.0, 0.0,
          0.0, 1.0, 1.0, 2.0, 1.0, 1.0, 0.0, 0.0, 1.0, 1.0, 3.0, 2.0, 0.0, 0.0, 0.0, 1.0, 2.0, 4.0]
    M6 = np.array(L6).reshape(6, 6, order="F")
    kats["lobpcg_6x6"] = {
        "source": "test/lobpcg_test.hpp:10-55",
        "n": 6, "input_colmajor": L6, "expected_top2_reference": [7.3191, 5.6639],
        "eigenvalues": np.linalg.eigvalsh(M6).tolist(), "tol": 1e-4,
    }
    kats["eigenpairs_5x5"] = {
        "source": "test/lanczos_test.hpp:44-84",
        "n": 5, "input_colmajor": L5, "k": 3, "two_norm_reference": 7.0890,
        "eigenvalues": np.linalg.eigvalsh(M5).tolist(),
    }
    kats["simple5_deflate"] = {
        "source": "test/composite_FP32_test.hpp:249-290",
        "n": 5, "diagonal": d5, "expected_diagonal": [max(x, 0.0) for x in d5], "k": 2, "tol": 1e-3,
        "entry": "composite_FP32_auto_scale_deflate",
    }
    (HERE / "reference_kats.json").write_text(json.dumps(kats, indent=1))

    x = np.linspace(-1.0, 1.0, 4001)
    np.savez_compressed(HERE / "transfer_fp32.npz", x=x,
                        g=O.transfer_function(x, O.fp32_chain_coefficients()),
                        coeffs=np.array(O.fp32_chain_coefficients(), dtype=np.float32))
    np.savez_compressed(HERE / "transfer_fp16.npz", x=x,
                        g=O.transfer_function(x, O.fp16_chain_coefficients()),
                        coeffs=np.array(O.fp16_chain_coefficients(), dtype=np.float32))

    out = {}
    for n, seed in ((64, 11), (128, 12)):
        X, Q, lam = O.spectral_input(n, seed, "uniform")
        out[f"X{n}"] = X
        out[f"Q{n}"] = Q
        out[f"lam{n}"] = lam
        out[f"chain_fp32_table_f64_{n}"] = O.composite_exact(X, "fp32")
        out[f"chain_fp16_table_f64_{n}"] = O.composite_exact(X, "fp16")
    np.savez_compressed(HERE / "chain_small.npz", **out)
    print("wrote", sorted(p.name for p in HERE.iterdir()))


if __name__ == "__main__":
    main()
