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
compute_eigenpairs           -> compute_eigenpairs(handle, A, k, ...) -> (values, vectors, bound)
lobpcg                       -> lobpcg(handle, A, m, ...) -> (V, D)
composite_FP32_auto_scale_deflate, composite_FP16_auto_scale_deflate, eig_FP64_deflate

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
FMT_F32, FMT_F16, FMT_F16X3, FMT_BF16, FMT_BF16X3, FMT_BF16X9, FMT_E4M3, FMT_F16MX = range(8)
FORMATS = {"f32": FMT_F32, "f16": FMT_F16, "f16x3": FMT_F16X3, "bf16": FMT_BF16,
           "bf16x3": FMT_BF16X3, "bf16x9": FMT_BF16X9, "e4m3": FMT_E4M3, "f16mx": FMT_F16MX}
# psdp_table_t
TABLE_FP32, TABLE_FP16 = 0, 1
# psdp_method_t
METHOD_FP32, METHOD_FP16, METHOD_FP32_EMULATED, METHOD_MIXED, METHOD_CUSTOM, METHOD_MIXED_MX = range(6)
METHODS = {"composite_FP32": METHOD_FP32, "composite_FP16": METHOD_FP16,
           "composite_FP32_emulated": METHOD_FP32_EMULATED, "mixed": METHOD_MIXED,
           "custom": METHOD_CUSTOM, "mixed_mx": METHOD_MIXED_MX}
# The MEX gateway's method strings (MATLAB/psd_projection_MATLAB.cu:112-166,
# README.md:56) plus the shipped schedule; psdp_project_method dispatches them.
MEX_METHODS = ("composite_FP16", "composite_FP32", "composite_FP32_emulated", "eig_FP64", "eig_FP32", "mixed",
               "mixed_mx")
EVD_METHODS = ("eig_FP64", "eig_FP32")
# psdp_bound_mode_t
BOUND_REFERENCE, BOUND_TIGHT = 0, 1
MAX_STEPS = 16
# psdp_prof_kind_t
PROF_KINDS = {0: "prod_f32", 1: "prod_f16", 2: "prod_f16x3", 3: "prod_bf16", 4: "prod_bf16x3",
              5: "prod_bf16x9", 6: "prod_e4m3", 7: "prod_f16mx", 8: "lanczos", 9: "prep", 10: "eigenpairs", 11: "lobpcg",
              12: "rank_update", 13: "graph"}
# psdp_ritz_residual_t
RITZ_RESIDUAL_FIRST_ROW, RITZ_RESIDUAL_LAST_ROW = 0, 1


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
        ("bound_mode", ctypes.c_int),
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
    "psdp_last_bounds_batched": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                                                ctypes.c_int]),
    "psdp_eig_fp64_psd_alloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                               ctypes.POINTER(ctypes.c_void_p)]),
    "psdp_eig_fp32_psd_alloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                               ctypes.POINTER(ctypes.c_void_p)]),
    "psdp_free": (ctypes.c_int, [ctypes.c_void_p]),
    "psdp_project_method": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int64,
                                           ctypes.c_void_p]),
    "psdp_approximate_two_norm": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                 ctypes.POINTER(ctypes.c_double),
                                                 ctypes.POINTER(ctypes.c_double), ctypes.c_int64,
                                                 ctypes.c_double]),
    "psdp_project_f64": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.POINTER(Config), ctypes.c_void_p, ctypes.c_size_t]),
    "psdp_project_f32": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.POINTER(Config), ctypes.c_void_p, ctypes.c_size_t]),
    "psdp_project_f32_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                ctypes.POINTER(Config)]),
    "psdp_session_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Config), ctypes.c_size_t,
                                           ctypes.POINTER(ctypes.c_void_p)]),
    "psdp_session_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "psdp_session_block_bytes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_size_t)]),
    "psdp_session_add": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.POINTER(ctypes.c_int64)]),
    "psdp_session_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "psdp_session_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]),
    "psdp_sym_product_f32": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int]),
    "psdp_project_batched_f32": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Block), ctypes.c_int,
                                                ctypes.POINTER(Config)]),
    "psdp_project_packed_f32": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.POINTER(ctypes.c_int64), ctypes.c_int, ctypes.POINTER(Config)]),
    "psdp_workspace_size_batched": (ctypes.c_int, [ctypes.POINTER(Block), ctypes.c_int, ctypes.POINTER(Config),
                                                   ctypes.POINTER(ctypes.c_size_t)]),
    "psdp_profile_enable": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "psdp_profile_summary": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                            ctypes.POINTER(ctypes.c_double)]),
    "psdp_eig_fp64_psd": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
    "psdp_eig_fp32_psd": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
    "psdp_set_eigenpair_options": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "psdp_set_batched_tiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "psdp_set_split_k": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "psdp_set_batched_graph": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "psdp_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(Config)]),
    "psdp_tuned_tile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int)]),
    "psdp_set_tuned_tile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "psdp_autotune_batched": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Block), ctypes.c_int,
                                             ctypes.POINTER(Config)]),
    "psdp_compute_eigenpairs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                               ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                               ctypes.POINTER(ctypes.c_double)]),
    "psdp_lobpcg": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
    "psdp_lobpcg_last": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_double)]),
    "psdp_composite_fp32_auto_scale_deflate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                              ctypes.c_int64, ctypes.c_double, ctypes.c_double]),
    "psdp_composite_fp16_auto_scale_deflate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                              ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int64,
                                                              ctypes.c_int, ctypes.c_double]),
    "psdp_eig_fp64_deflate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                             ctypes.c_int, ctypes.c_double]),
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
        elif k == "bound_mode":
            cfg.bound_mode = {"reference": BOUND_REFERENCE, "tight": BOUND_TIGHT}.get(v, v)
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

    def last_bounds_batched(self, count):
        """[(lo, up, iterations)] of each block of the last batched call."""
        lo = (ctypes.c_double * count)()
        up = (ctypes.c_double * count)()
        it = (ctypes.c_int * count)()
        _check(lib().psdp_last_bounds_batched(self._h, lo, up, it, int(count)))
        return [(lo[i], up[i], it[i]) for i in range(count)]

    def set_eigenpair_options(self, residual_row="first", normalize=False):
        """compute_eigenpairs options: residual row "first" (the reference's)
        or "last"; normalize divides the kept Ritz vectors by their norms."""
        m = {"first": RITZ_RESIDUAL_FIRST_ROW, "last": RITZ_RESIDUAL_LAST_ROW}.get(residual_row, residual_row)
        _check(lib().psdp_set_eigenpair_options(self._h, int(m), 1 if normalize else 0))

    def set_split_k(self, on=True):
        """Split products (default on): a last round of 256 x 256 tiles that
        fills at most half the CUs runs its tiles' K range in chunks side by
        side; off, it is re-cut into 64 x 64 tiles and every tile size gives
        bitwise the same result."""
        _check(lib().psdp_set_split_k(self._h, 1 if on else 0))

    def set_batched_tiling(self, big_block_n=0, concurrent=False):
        """Blocks with n >= big_block_n run on 128 x 128 tiles in batched calls
        (0: default); concurrent runs the two tile classes on two streams."""
        _check(lib().psdp_set_batched_tiling(self._h, int(big_block_n), 1 if concurrent else 0))

    def set_batched_graph(self, on=True):
        """Capture the device-side sequence of batched calls in a HIP graph and
        replay it for calls with the same plan."""
        _check(lib().psdp_set_batched_graph(self._h, 1 if on else 0))

    def autotune(self, n, method="mixed", config: Config | None = None, **overrides):
        """Times the product tile sizes for projections of size n and keeps the
        fastest; returns {(format, averaged): tile edge}."""
        cfg = config if config is not None else default_config(method, **overrides)
        _check(lib().psdp_autotune(self._h, int(n), ctypes.byref(cfg)))
        out = {}
        fmts = [cfg.step_format[i] for i in range(MAX_STEPS)] + [cfg.final_format]
        for f in sorted(set(FORMATS.values())):
            for avg in (0, 1):
                t = ctypes.c_int()
                _check(lib().psdp_tuned_tile(self._h, int(n), f, avg, ctypes.byref(t)))
                if t.value:
                    out[(f, avg)] = t.value
        return out

    def set_tile(self, n, fmt, tile, average=False):
        """Tile edge (64/128/256, 0 clears) of products of size n in format fmt."""
        f = FORMATS[fmt] if isinstance(fmt, str) else int(fmt)
        _check(lib().psdp_set_tuned_tile(self._h, int(n), f, 1 if average else 0, int(tile)))

    def autotune_batched(self, blocks, outs, method="mixed", config: Config | None = None, **overrides):
        """Picks the batched tile classes by timing these blocks (writes outs)."""
        self._sync_stream()
        arr = (Block * len(blocks))()
        for i, (x, y) in enumerate(zip(blocks, outs)):
            arr[i].X = _ptr(x)
            arr[i].Y = _ptr(y)
            arr[i].n = _size(x, None)
        cfg = config if config is not None else default_config(method, **overrides)
        _check(lib().psdp_autotune_batched(self._h, arr, len(blocks), ctypes.byref(cfg)))

    def lobpcg_last(self):
        """(iterations, ||R||_F) of the last lobpcg run."""
        it, rn = ctypes.c_int(), ctypes.c_double()
        _check(lib().psdp_lobpcg_last(self._h, ctypes.byref(it), ctypes.byref(rn)))
        return it.value, rn.value

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


class Session:
    """psdp_session_t: continuous batching of float32 blocks on a handle's
    stream.  add() admits a block (None when the arena has no room now);
    step() enqueues the prologue of new blocks and one product of every ready
    block, returns the ticks in flight; poll() returns (completed tickets,
    blocks not yet complete).  A block's result equals project_batched alone."""

    def __init__(self, handle, arena_bytes, method="mixed", config: Config | None = None, **overrides):
        self.handle = handle
        cfg = config if config is not None else default_config(method, **overrides)
        self._s = ctypes.c_void_p()
        _check(lib().psdp_session_create(handle.ptr, ctypes.byref(cfg), int(arena_bytes), ctypes.byref(self._s)))
        self._ids = (ctypes.c_int64 * 4096)()

    def block_bytes(self, n):
        b = ctypes.c_size_t()
        _check(lib().psdp_session_block_bytes(self._s, int(n), ctypes.byref(b)))
        return b.value

    def add(self, X, Y):
        n = _size(X, None)
        t = ctypes.c_int64()
        st = lib().psdp_session_add(self._s, _ptr(X), _ptr(Y), n, ctypes.byref(t))
        if st == ERR_NOMEM:
            return None
        _check(st)
        return t.value

    def step(self):
        k = ctypes.c_int()
        _check(lib().psdp_session_step(self._s, ctypes.byref(k)))
        return k.value

    def poll(self):
        """(completed tickets, blocks not yet complete, ticks in flight)."""
        c, a, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        out = []
        while True:
            _check(lib().psdp_session_poll(self._s, self._ids, len(self._ids), ctypes.byref(c), ctypes.byref(a),
                                           ctypes.byref(k)))
            out.extend(self._ids[i] for i in range(c.value))
            if c.value < len(self._ids):
                return out, a.value, k.value

    def close(self):
        if self._s:
            lib().psdp_session_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


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


def eig_psd_alloc(handle, mat, n=None, fp64=True, return_eigenvalues=True):
    """The reference's ownership of eig_FP64_psd / eig_FP32_psd: returns the
    device address of n eigenvalues the library allocated (the caller frees
    it with free_device), or 0 when not requested."""
    handle._sync_stream()
    n = _size(mat, n)
    out = ctypes.c_void_p()
    fn = lib().psdp_eig_fp64_psd_alloc if fp64 else lib().psdp_eig_fp32_psd_alloc
    _check(fn(handle.ptr, _ptr(mat), n, ctypes.byref(out) if return_eigenvalues else None))
    return out.value or 0


def free_device(ptr):
    """psdp_free: releases device memory the library allocated for the caller."""
    _check(lib().psdp_free(ctypes.c_void_p(ptr) if ptr else None))


def project_method(handle, mat, method, n=None, return_eigenvalues=False):
    """The MEX gateway's dispatch (MATLAB/psd_projection_MATLAB.cu:112-166):
    projects the float64 matrix in place by method name (composite_FP16,
    composite_FP32, composite_FP32_emulated, eig_FP64, eig_FP32, mixed).
    Returns (mat, eigenvalues) -- eigenvalues only for eig_* with
    return_eigenvalues, else None."""
    import torch
    handle._sync_stream()
    n = _size(mat, n)
    if mat.dtype != torch.float64:
        raise TypeError("expected a float64 matrix")
    w = None
    if return_eigenvalues and method in EVD_METHODS:
        w = torch.empty(n, dtype=torch.float64, device=mat.device)
    _check(lib().psdp_project_method(handle.ptr, str(method).encode(), _ptr(mat), n, _ptr(w)))
    return mat, w


def _f64_square(A, n):
    import torch
    n = _size(A, n)
    if A.dtype != torch.float64:
        raise TypeError("expected a float64 matrix")
    return n


def compute_eigenpairs(handle, A, k, n=None, upper_bound_only=False, max_iter=0, tol=1e-10, ortho_tol=1e-5):
    """src/lanczos.cu:319-547 (lanczos.h:33-43 defaults).  Returns
    (eigenvalues[:r], eigenvectors[:, :r], bound); eigenvectors are the
    unnormalised Ritz vectors as columns.  With upper_bound_only, (None, None,
    bound)."""
    import torch
    handle._sync_stream()
    n = _f64_square(A, n)
    b = ctypes.c_double()
    if upper_bound_only:
        _check(lib().psdp_compute_eigenpairs(handle.ptr, _ptr(A), n, int(k), None, None, None, 1, int(max_iter),
                                             float(tol), float(ortho_tol), ctypes.byref(b)))
        return None, None, b.value
    kk = max(int(k), 1)
    vals = torch.empty(kk, dtype=torch.float64, device=A.device)
    vecs = torch.empty((kk, n), dtype=torch.float64, device=A.device)   # row i = column i of the n x k array
    r = ctypes.c_int64()
    _check(lib().psdp_compute_eigenpairs(handle.ptr, _ptr(A), n, int(k), ctypes.byref(r), _ptr(vals), _ptr(vecs),
                                         0, int(max_iter), float(tol), float(ortho_tol), ctypes.byref(b)))
    return vals[:r.value], vecs[:r.value].T, b.value


def lobpcg(handle, A, m, n=None, V=None, D=None, warmstart=False, maxiter=100, tol=1e-8):
    """src/lobpcg.cu:12-256 (lobpcg.h:16-28 defaults): the m largest
    eigenpairs.  Returns (V, D): V is n x m with the eigenvectors as columns
    (column-major storage, i.e. V.T is a contiguous m x n tensor), D the m
    eigenvalues in descending order.  With warmstart, V (a tensor of that
    layout) is the orthonormal start block."""
    import torch
    handle._sync_stream()
    n = _f64_square(A, n)
    if V is None:
        Vt = torch.empty((m, n), dtype=torch.float64, device=A.device)
    else:
        Vt = V.T
        if tuple(Vt.shape) != (m, n) or not Vt.is_contiguous():
            raise ValueError("V must be an n x m column-major tensor (V.T contiguous)")
    if D is None:
        D = torch.empty(m, dtype=torch.float64, device=A.device)
    _check(lib().psdp_lobpcg(handle.ptr, _ptr(A), _ptr(Vt), _ptr(D), n, int(m), 1 if warmstart else 0,
                             int(maxiter), float(tol)))
    return Vt.T, D


def composite_FP32_auto_scale_deflate(handle, mat, n=None, k=40, tol=1e-10, ortho_tol=1e-3):
    """src/composite_FP32.cu:193-254 (composite_FP32.h:36-45 defaults)."""
    handle._sync_stream()
    n = _f64_square(mat, n)
    _check(lib().psdp_composite_fp32_auto_scale_deflate(handle.ptr, _ptr(mat), n, int(k), float(tol),
                                                        float(ortho_tol)))
    return mat


def composite_FP16_auto_scale_deflate(handle, mat, n=None, workspace=None, k=40, maxiter=100, tol=1e-10):
    """src/composite_FP16.cu:276-380 (composite_FP16.h:39-50 defaults)."""
    handle._sync_stream()
    n = _f64_square(mat, n)
    ws, wsb = _ws(workspace)
    _check(lib().psdp_composite_fp16_auto_scale_deflate(handle.ptr, _ptr(mat), n, ws, wsb, int(k), int(maxiter),
                                                        float(tol)))
    return mat


def eig_FP64_deflate(handle, mat, n=None, k=40, maxiter=100, tol=1e-10):
    """src/eig_FP64_psd.cu:87-177 (eig_FP64_psd.h:22-31 defaults)."""
    handle._sync_stream()
    n = _f64_square(mat, n)
    _check(lib().psdp_eig_fp64_deflate(handle.ptr, _ptr(mat), n, int(k), int(maxiter), float(tol)))
    return mat


def project(handle, X, out=None, method="mixed", config: Config | None = None, workspace=None, **overrides):
    """Y = Pi(X) for a float32 or float64 symmetric matrix (X may be `out`)."""
    import torch
    handle._sync_stream()
    n = _size(X, None)
    if out is None:
        out = torch.empty_like(X)
    if config is None and method in EVD_METHODS:
        # the EVD methods of the MEX set: float64 in place on a copy
        tmp = X.to(torch.float64, copy=True)
        project_method(handle, tmp, method)
        out.copy_(tmp)
        return out
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


def project_noalloc(handle, X, out=None, method="mixed", config: Config | None = None, **overrides):
    """psdp_project_f32_noalloc: Y = Pi(X) for a float32 matrix with no device
    allocation in the call (workspace from the library's per-GPU scratch
    array; calls that overlap on one GPU race)."""
    import torch
    handle._sync_stream()
    n = _size(X, None)
    if X.dtype != torch.float32:
        raise TypeError("expected a float32 matrix")
    if out is None:
        out = torch.empty_like(X)
    cfg = config if config is not None else default_config(method, **overrides)
    _check(lib().psdp_project_f32_noalloc(handle.ptr, _ptr(X), _ptr(out), n, ctypes.byref(cfg)))
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


def project_packed(handle, data, sizes, out=None, method="mixed", config: Config | None = None, **overrides):
    """Projects a packed block-diagonal cone: `data` is a 1-D float32 device
    tensor holding column-major blocks of the given sizes back to back
    (block b starts at sum of sizes[c]**2 for c < b).  Writes `out` (default:
    `data`, in place) in the same layout and returns it."""
    sizes = [int(v) for v in sizes]
    if data.dim() != 1 or not data.is_contiguous():
        raise ValueError("packed data must be a contiguous 1-D tensor")
    if sum(v * v for v in sizes) != data.numel():
        raise ValueError(f"sizes cover {sum(v * v for v in sizes)} elements, data holds {data.numel()}")
    if out is None:
        out = data
    elif out.shape != data.shape or not out.is_contiguous():
        raise ValueError("out must be a contiguous tensor shaped like data")
    handle._sync_stream()
    arr = (ctypes.c_int64 * len(sizes))(*sizes)
    cfg = config if config is not None else default_config(method, **overrides)
    _check(lib().psdp_project_packed_f32(handle.ptr, _ptr(data), _ptr(out), arr, len(sizes), ctypes.byref(cfg)))
    return out


__all__ = [
    "Handle", "PsdpError", "Config", "default_config", "lib",
    "composite_FP32", "composite_FP32_auto_scale", "composite_FP16", "composite_FP16_auto_scale",
    "composite_FP32_emulated", "composite_FP32_emulated_auto_scale", "approximate_two_norm",
    "eig_FP64_psd", "eig_FP32_psd", "project", "project_batched", "project_packed", "sym_product", "workspace_size",
    "compute_eigenpairs", "lobpcg", "composite_FP32_auto_scale_deflate", "composite_FP16_auto_scale_deflate",
    "eig_FP64_deflate", "project_method", "eig_psd_alloc", "free_device", "project_noalloc", "Session",
]
