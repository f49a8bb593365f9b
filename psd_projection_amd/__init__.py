"""psd_projection_amd: PSD projection on AMD MI355X (gfx950).

Python mirror of the reference's entry points (ComputationalRobotics/
psd_projection, include/psd_projection/*.h) over the C ABI of libpsdp.so
(include/psd_projection/psdp.h).  Matrices are CUDA/HIP torch tensors (or raw
device pointers); every call runs the HIP kernels of libpsdp.  There is no CPU
fallback: if the library is missing, importing a function that needs it
raises.

Reference name               -> here
composite_FP32               -> composite_FP32(handle, mat, n=None, workspace=None)
composite_FP32_auto_scale    -> composite_FP32_auto_scale(handle, mat, ...)
composite_FP16(_auto_scale)  -> composite_FP16(_auto_scale)(handle, mat, ...)
composite_FP32_emulated(_auto_scale)
approximate_two_norm         -> approximate_two_norm(handle, A, n=None, max_iter=20, tol=1e-10) -> (lo, up)
eig_FP64_psd / eig_FP32_psd  -> eig_FP64_psd(handle, mat, n=None, return_eigenvalues=False)

The reference passes separate cuBLAS and cuSOLVER handles; here one Handle
(device + stream + workspace arena) takes their place.  The reference prints
errors and continues; here a failing call raises PsdpError carrying the
library's status and message.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libpsdp.so"

# psdp_status_t
OK, ERR_INVALID_ARG, ERR_NOMEM, ERR_NONFINITE, ERR_UNSUPPORTED, ERR_DEVICE = range(6)
# psdp_format_t
FMT_F32, FMT_F16, FMT_F16X3, FMT_BF16, FMT_BF16X3, FMT_BF16X9, FMT_E4M3 = range(7)
FORMATS = {"f32": FMT_F32, "f16": FMT_F16, "f16x3": FMT_F16X3, "bf16": FMT_BF16,
           "bf16x3": FMT_BF16X3, "bf16x9": FMT_BF16X9, "e4m3": FMT_E4M3}
# psdp_table_t
TABLE_FP32, TABLE_FP16 = 0, 1
# psdp_method_t
METHOD_FP32, METHOD_FP16, METHOD_FP32_EMULATED, METHOD_MIXED, METHOD_CUSTOM = range(5)
METHODS = {"composite_FP32": METHOD_FP32, "composite_FP16": METHOD_FP16,
           "composite_FP32_emulated": METHOD_FP32_EMULATED, "mixed": METHOD_MIXED,
           "custom": METHOD_CUSTOM}
MAX_STEPS = 16
# psdp_prof_kind_t
PROF_KINDS = {0: "prod_f32", 1: "prod_f16", 2: "prod_f16x3", 3: "prod_bf16", 4: "prod_bf16x3",
              5: "prod_bf16x9", 6: "prod_e4m3", 8: "lanczos", 9: "prep"}


class PsdpError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{_STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


_STATUS_NAMES = {0: "PSDP_OK", 1: "PSDP_ERR_INVALID_ARG", 2: "PSDP_ERR_NOMEM",
                 3: "PSDP_ERR_NONFINITE", 4: "PSDP_ERR_UNSUPPORTED", 5: "PSDP_ERR_DEVICE"}


class Config(ctypes.Structure):
    """psdp_config_t."""
    _fields_ = [
        ("method", ctypes.c_int),
        ("table", ctypes.c_int),
        ("step_format", ctypes.c_int * MAX_STEPS),
        ("final_format", ctypes.c_int),
        ("auto_scale", ctypes.c_int),
        ("lanczos_iters", ctypes.c_int),
        ("lanczos_tol", ctypes.c_double),
        ("scale_margin", ctypes.c_double),
        ("synchronize", ctypes.c_int),
        ("average_products", ctypes.c_int),
    ]


class Block(ctypes.Structure):
    """psdp_block_t."""
    _fields_ = [("X", ctypes.c_void_p), ("Y", ctypes.c_void_p), ("n", ctypes.c_int64)]


_lib_handle = None

_PROTOS = {
    "psdp_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]),
    "psdp_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "psdp_set_stream": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "psdp_get_stream": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "psdp_set_seed": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64]),
    "psdp_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "psdp_last_error": (ctypes.c_char_p, []),
    "psdp_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "psdp_default_config": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.c_int]),
    "psdp_workspace_size": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.c_int64,
                                           ctypes.POINTER(ctypes.c_size_t)]),
    "psdp_last_bound": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]),
    "psdp_approximate_two_norm": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                 ctypes.POINTER(ctypes.c_double),
                                                 ctypes.POINTER(ctypes.c_double), ctypes.c_int64,
                                                 ctypes.c_double]),
    "psdp_project_f64": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.POINTER(Config), ctypes.c_void_p, ctypes.c_size_t]),
    "psdp_project_f32": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.POINTER(Config), ctypes.c_void_p, ctypes.c_size_t]),
    "psdp_sym_product_f32": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int]),
    "psdp_project_batched_f32": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Block), ctypes.c_int,
                                                ctypes.POINTER(Config)]),
    "psdp_workspace_size_batched": (ctypes.c_int, [ctypes.POINTER(Block), ctypes.c_int, ctypes.POINTER(Config),
                                                   ctypes.POINTER(ctypes.c_size_t)]),
    "psdp_profile_enable": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "psdp_profile_summary": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                            ctypes.POINTER(ctypes.c_double)]),
    "psdp_eig_fp64_psd": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
    "psdp_eig_fp32_psd": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
}
for _name in ("fp32", "fp32_auto_scale", "fp16", "fp16_auto_scale", "fp32_emulated",
              "fp32_emulated_auto_scale"):
    _PROTOS[f"psdp_composite_{_name}"] = (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                          ctypes.c_void_p, ctypes.c_size_t])


def lib():
    """Loads libpsdp.so (built in-tree by psd_projection_amd._build)."""
    global _lib_handle
    if _lib_handle is None:
        path = os.environ.get("PSDP_LIB", str(LIB_PATH))
        if not Path(path).exists():
            raise ImportError(f"libpsdp.so not found at {path}; build it with "
                              "`python -m psd_projection_amd._build` (hipcc, gfx950)")
        L = ctypes.CDLL(path)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib_handle = L
    return _lib_handle


def _check(status):
    if status != OK:
        msg = lib().psdp_last_error().decode(errors="replace")
        raise PsdpError(status, msg)


def default_config(method="mixed", **overrides) -> Config:
    cfg = Config()
    m = METHODS[method] if isinstance(method, str) else int(method)
    _check(lib().psdp_default_config(ctypes.byref(cfg), m))
    for k, v in overrides.items():
        if k == "step_format":
            for i, f in enumerate(v):
                cfg.step_format[i] = FORMATS[f] if isinstance(f, str) else int(f)
        elif k == "final_format":
            cfg.final_format = FORMATS[v] if isinstance(v, str) else int(v)
        elif k == "table":
            cfg.table = {"fp32": TABLE_FP32, "fp16": TABLE_FP16}.get(v, v)
        else:
            setattr(cfg, k, v)
    return cfg


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _size(mat, n):
    if n is not None:
        return int(n)
    shape = tuple(mat.shape)
    if len(shape) != 2 or shape[0] != shape[1]:
        raise ValueError(f"expected a square matrix, got shape {shape}")
    if not mat.is_contiguous():
        raise ValueError("matrix must be contiguous")
    return shape[0]


class Handle:
    """psdp_handle_t: device, stream, workspace arena, Lanczos state."""

    def __init__(self, device: int = 0, stream=None, seed: int | None = None):
        self._h = ctypes.c_void_p()
        _check(lib().psdp_create(ctypes.byref(self._h), int(device)))
        self.device = int(device)
        self._follow_torch = stream is None
        if stream is not None:
            self.set_stream(stream)
        if seed is not None:
            _check(lib().psdp_set_seed(self._h, int(seed)))

    @property
    def ptr(self):
        return self._h

    def set_stream(self, stream):
        s = stream if isinstance(stream, int) else getattr(stream, "cuda_stream", stream)
        _check(lib().psdp_set_stream(self._h, ctypes.c_void_p(s)))
        self._follow_torch = False

    def _sync_stream(self):
        # Run on torch's current stream so that torch ops before and after the
        # call are ordered with it.
        if self._follow_torch:
            import torch
            s = torch.cuda.current_stream(self.device).cuda_stream
            _check(lib().psdp_set_stream(self._h, ctypes.c_void_p(s)))
            self._follow_torch = True

    def set_seed(self, seed):
        _check(lib().psdp_set_seed(self._h, int(seed)))

    def synchronize(self):
        _check(lib().psdp_synchronize(self._h))

    def last_bound(self):
        lo, up, it = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        _check(lib().psdp_last_bound(self._h, ctypes.byref(lo), ctypes.byref(up), ctypes.byref(it)))
        return lo.value, up.value, it.value

    def profile(self, on=True):
        """Start (on=True, clears old records) or stop event timing of launches."""
        _check(lib().psdp_profile_enable(self._h, 1 if on else 0))

    def profile_summary(self):
        """{kind: (launches, total_ms)} of the recorded launches."""
        out = {}
        for k, name in PROF_KINDS.items():
            c, t = ctypes.c_int(), ctypes.c_double()
            _check(lib().psdp_profile_summary(self._h, k, ctypes.byref(c), ctypes.byref(t)))
            if c.value:
                out[name] = (c.value, t.value)
        return out

    def close(self):
        if self._h:
            lib().psdp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _ws(workspace):
    if workspace is None:
        return None, 0
    return _ptr(workspace), int(workspace.numel() * workspace.element_size())


def _dropin(cname, handle, mat, n, workspace):
    handle._sync_stream()
    n = _size(mat, n)
    ws, wsb = _ws(workspace)
    _check(getattr(lib(), f"psdp_composite_{cname}")(handle.ptr, _ptr(mat), n, ws, wsb))
    return mat


def composite_FP32(handle, mat, n=None, workspace=None):
    """src/composite_FP32.cu:15-164: in-place projection of an already scaled
    (spectrum in [-1,1]) symmetric FP64 matrix, FP32 chain."""
    return _dropin("fp32", handle, mat, n, workspace)


def composite_FP32_auto_scale(handle, mat, n=None, workspace=None):
    """src/composite_FP32.cu:166-191."""
    return _dropin("fp32_auto_scale", handle, mat, n, workspace)


def composite_FP16(handle, mat, n=None, workspace=None):
    """src/composite_FP16.cu:16-246."""
    return _dropin("fp16", handle, mat, n, workspace)


def composite_FP16_auto_scale(handle, mat, n=None, workspace=None):
    """src/composite_FP16.cu:248-274."""
    return _dropin("fp16_auto_scale", handle, mat, n, workspace)


def composite_FP32_emulated(handle, mat, n=None, workspace=None):
    """src/composite_FP32_emulated.cu:20-235 (BF16x9 split products)."""
    return _dropin("fp32_emulated", handle, mat, n, workspace)


def composite_FP32_emulated_auto_scale(handle, mat, n=None, workspace=None):
    """src/composite_FP32_emulated.cu:237-262."""
    return _dropin("fp32_emulated_auto_scale", handle, mat, n, workspace)


def approximate_two_norm(handle, A, n=None, max_iter=20, tol=1e-10):
    """src/lanczos.cu:41-293 -> (lo, up)."""
    handle._sync_stream()
    n = _size(A, n)
    lo, up = ctypes.c_double(), ctypes.c_double()
    _check(lib().psdp_approximate_two_norm(handle.ptr, _ptr(A), n, ctypes.byref(lo), ctypes.byref(up),
                                           int(max_iter), float(tol)))
    return lo.value, up.value


def eig_FP64_psd(handle, mat, n=None, return_eigenvalues=False):
    """src/eig_FP64_psd.cu:24-85 (rocSOLVER dsyevd).  Returns the eigenvalue
    tensor when requested (the reference returns a device pointer)."""
    handle._sync_stream()
    n = _size(mat, n)
    w = None
    if return_eigenvalues:
        import torch
        w = torch.empty(n, dtype=torch.float64, device=mat.device)
    _check(lib().psdp_eig_fp64_psd(handle.ptr, _ptr(mat), n, _ptr(w)))
    return w


def eig_FP32_psd(handle, mat, n=None, return_eigenvalues=False):
    """src/eig_FP32_psd.cu:23-96 (rocSOLVER ssyevd)."""
    handle._sync_stream()
    n = _size(mat, n)
    w = None
    if return_eigenvalues:
        import torch
        w = torch.empty(n, dtype=torch.float64, device=mat.device)
    _check(lib().psdp_eig_fp32_psd(handle.ptr, _ptr(mat), n, _ptr(w)))
    return w


def project(handle, X, out=None, method="mixed", config: Config | None = None, workspace=None, **overrides):
    """Y = Pi(X) for a float32 or float64 symmetric matrix (X may be `out`)."""
    import torch
    handle._sync_stream()
    n = _size(X, None)
    if out is None:
        out = torch.empty_like(X)
    cfg = config if config is not None else default_config(method, **overrides)
    ws, wsb = _ws(workspace)
    if X.dtype == torch.float64:
        fn = lib().psdp_project_f64
    elif X.dtype == torch.float32:
        fn = lib().psdp_project_f32
    else:
        raise TypeError(f"unsupported dtype {X.dtype}")
    if out.dtype != X.dtype:
        raise TypeError("out must have the dtype of X")
    _check(fn(handle.ptr, _ptr(X), _ptr(out), n, ctypes.byref(cfg), ws, wsb))
    return out


def workspace_size(n, method="mixed", **overrides) -> int:
    cfg = default_config(method, **overrides)
    b = ctypes.c_size_t()
    _check(lib().psdp_workspace_size(ctypes.byref(cfg), int(n), ctypes.byref(b)))
    return b.value


def sym_product(handle, X, Y, fmt="f32", out=None, average=False):
    """Symmetric product of two symmetric float32 matrices with operands in
    `fmt`: the upper triangle of X Y mirrored, or (X Y + Y X)/2 with average."""
    import torch
    handle._sync_stream()
    n = _size(X, None)
    if out is None:
        out = torch.empty_like(X)
    f = FORMATS[fmt] if isinstance(fmt, str) else int(fmt)
    _check(lib().psdp_sym_product_f32(handle.ptr, _ptr(X), _ptr(Y), _ptr(out), n, f, 1 if average else 0))
    return out


def workspace_size_batched(sizes, method="mixed", **overrides) -> int:
    """Workspace bytes of a batched call on blocks of the given sizes (host
    only; builds and checks the launch descriptors with stand-in pointers)."""
    arr = (Block * len(sizes))()
    for i, n in enumerate(sizes):
        arr[i].X = 0x1000 + 16 * i
        arr[i].Y = 0x2000 + 16 * i
        arr[i].n = int(n)
    cfg = default_config(method, **overrides)
    b = ctypes.c_size_t()
    _check(lib().psdp_workspace_size_batched(arr, len(sizes), ctypes.byref(cfg), ctypes.byref(b)))
    return b.value


def project_batched(handle, blocks, outs=None, method="mixed", config: Config | None = None, **overrides):
    """Projects a list of float32 symmetric blocks (one grouped launch per
    filter-chain product).  Returns the list of outputs."""
    import torch
    handle._sync_stream()
    if outs is None:
        outs = [torch.empty_like(b) for b in blocks]
    arr = (Block * len(blocks))()
    for i, (x, y) in enumerate(zip(blocks, outs)):
        arr[i].X = _ptr(x)
        arr[i].Y = _ptr(y)
        arr[i].n = _size(x, None)
    cfg = config if config is not None else default_config(method, **overrides)
    _check(lib().psdp_project_batched_f32(handle.ptr, arr, len(blocks), ctypes.byref(cfg)))
    return outs


__all__ = [
    "Handle", "PsdpError", "Config", "default_config", "lib",
    "composite_FP32", "composite_FP32_auto_scale", "composite_FP16", "composite_FP16_auto_scale",
    "composite_FP32_emulated", "composite_FP32_emulated_auto_scale", "approximate_two_norm",
    "eig_FP64_psd", "eig_FP32_psd", "project", "project_batched", "sym_product", "workspace_size",
]
