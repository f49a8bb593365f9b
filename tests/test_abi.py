"""The C ABI library loads and exports every entry point its headers declare;
host-only calls (configuration, workspace sizing, argument checks) work
without a GPU.  CPU only."""
import ctypes
import re
from pathlib import Path

import pytest

import psd_projection_amd as P

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").rglob("*.h"))


def declared_symbols():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        body = text[text.find('extern "C"'):] if 'extern "C"' in text else ""
        for m in re.finditer(r"\b(psdp_[a-z0-9_]+)\s*\(", body):
            names.add(m.group(1))
    return sorted(names)


def test_library_built():
    assert P.LIB_PATH.exists(), "run psd_projection_amd/_build.py (or __graft_entry__.build())"


@pytest.mark.parametrize("name", declared_symbols())
def test_symbol_exported(name):
    L = ctypes.CDLL(str(P.LIB_PATH))
    assert hasattr(L, name), f"{name} declared in include/ but not exported"


def test_declarations_found():
    names = declared_symbols()
    for must in ("psdp_composite_fp32_auto_scale", "psdp_composite_fp16_auto_scale",
                 "psdp_composite_fp32_emulated_auto_scale", "psdp_approximate_two_norm",
                 "psdp_eig_fp64_psd", "psdp_project_f32", "psdp_project_batched_f32"):
        assert must in names


def test_default_config_and_workspace():
    cfg = P.default_config("mixed")
    assert cfg.auto_scale == 1 and cfg.lanczos_iters == 20 and cfg.lanczos_tol == 1e-10
    ws_mixed = P.workspace_size(16384, "mixed")
    ws_fp32 = P.workspace_size(16384, "composite_FP32")
    N = 16384 * (16384 + 64)  # rows padded by 64 elements
    # FP32 chain: A, A^3, A^2, X_s in float32 (+ Lanczos vectors)
    assert 16 * N <= ws_fp32 < 16 * N + (8 << 20)
    # mixed: A, A^3 float32 + 2 binary16 pieces for each of A, A^2, A^3, one for X_s
    assert 22 * N <= ws_mixed < 22 * N + (8 << 20)


def test_invalid_arguments_rejected_on_host():
    with pytest.raises(P.PsdpError) as e:
        P.workspace_size(0)
    assert e.value.status == P.ERR_INVALID_ARG
    with pytest.raises(P.PsdpError):
        P.workspace_size(100, "custom", table=7)
    with pytest.raises(P.PsdpError):
        P.workspace_size(100, "mixed", lanczos_iters=65)
    L = P.lib()
    assert L.psdp_status_string(3) == b"PSDP_ERR_NONFINITE"
    assert L.psdp_composite_fp32(None, None, 10, None, 0) == P.ERR_INVALID_ARG


def test_create_without_gpu_fails_cleanly():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(P.PsdpError):
        P.Handle(0)


@pytest.mark.parametrize("method", ["mixed", "composite_FP32", "composite_FP16", "composite_FP32_emulated"])
@pytest.mark.parametrize("sizes", [[1], [33], [32, 1024, 77, 300], list(range(1, 40, 3))])
def test_batched_plan_host_checks(method, sizes):
    # Builds every descriptor of a batched call on the host (stand-in device
    # pointers) and checks that no operand or output buffer is missing; the
    # workspace covers at least each block's own buffers.
    b = P.workspace_size_batched(sizes, method)
    single = sum(P.workspace_size(-(-n // 64) * 64 if n > 1024 else n, method) for n in sizes)
    assert b > 0
    assert b >= 0.5 * single


def test_batched_plan_rejects_bad_input():
    with pytest.raises(P.PsdpError):
        P.workspace_size_batched([], "mixed")
    with pytest.raises(P.PsdpError):
        P.workspace_size_batched([10, 0], "mixed")
