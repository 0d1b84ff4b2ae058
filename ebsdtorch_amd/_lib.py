"""ctypes binding of libebsd_amd.so (declared in include/ebsd_amd.h).

The library is mandatory: every compute entry of this package goes through
it, and there is no CPU or torch fallback. Loading fails loudly when the
shared object is missing or was built against a different ABI version.
"""

from __future__ import annotations

import ctypes
import os
import threading

import torch

from ._build import LIB_PATH

EBSD_ABI_VERSION = 3

EBSD_OK = 0
EBSD_ERR_NULL = -1
EBSD_ERR_SHAPE = -2
EBSD_ERR_DTYPE = -3
EBSD_ERR_TOPK = -4
EBSD_ERR_WORKSPACE = -5
EBSD_ERR_ALIGN = -6
EBSD_ERR_SCRATCH = -7
EBSD_ERR_INTERNAL = -8
EBSD_SCHEDULE_INFO = 11
SAMPLE_MODES = {"bilinear": 0, "nearest": 1, "bicubic": 2}
PADDING_MODES = {"zeros": 0, "border": 1, "reflection": 2}
EBSD_NOALLOC_SCRATCH_BYTES = 64 << 20
EBSD_MAX_TOPK = 16

DTYPE_F32 = 0
DTYPE_BF16 = 1
DTYPE_F16 = 2
DTYPE_FP8_E4M3 = 3

FLAG_ZERO_MEAN = 1
FLAG_UNIT_NORM = 2
FLAG_RENORM_DIRS = 4

TORCH_DTYPE = {
    DTYPE_F32: torch.float32,
    DTYPE_BF16: torch.bfloat16,
    DTYPE_F16: torch.float16,
    DTYPE_FP8_E4M3: torch.uint8,
}

_c_int = ctypes.c_int
_c_void_p = ctypes.c_void_p
_c_size_t = ctypes.c_size_t

# name -> (restype, argtypes); mirrors include/ebsd_amd.h
_SIGNATURES = {
    "ebsd_abi_version": (_c_int, []),
    "ebsd_status_string": (ctypes.c_char_p, [_c_int]),
    "ebsd_pad_multiple": (_c_int, [_c_int]),
    "ebsd_project": (
        _c_int,
        [_c_void_p, _c_int, _c_int, _c_void_p, _c_int, _c_void_p, _c_int, _c_void_p,
         _c_int, _c_int, _c_void_p, _c_int, _c_void_p],
    ),
    "ebsd_project_quads": (
        _c_int, [_c_void_p, _c_int, _c_int, _c_void_p, _c_int, _c_void_p, _c_int, _c_void_p, _c_int, _c_int,
                 _c_void_p, _c_int, _c_void_p, _c_void_p]),
    "ebsd_cmif_maps": (_c_int, [_c_void_p, ctypes.c_longlong, _c_int, _c_void_p, _c_void_p]),
    "ebsd_view_keys": (
        _c_int, [_c_void_p, _c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, _c_void_p, _c_void_p]),
    "ebsd_project_shifted": (
        _c_int,
        [_c_void_p, _c_int, _c_int, _c_void_p, _c_int, _c_void_p, _c_void_p, _c_int, _c_void_p,
         _c_int, _c_int, _c_void_p, _c_int, _c_void_p],
    ),
    "ebsd_poisson_gaussian_noise": (
        _c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_void_p, _c_int, _c_void_p, ctypes.c_float,
                 ctypes.c_ulonglong, ctypes.c_ulonglong, _c_void_p]),
    "ebsd_sample_dirs": (
        _c_int, [_c_void_p, _c_int, _c_int, _c_void_p, ctypes.c_longlong, _c_void_p, _c_int, _c_void_p]),
    "ebsd_sample_dirs_mode": (
        _c_int, [_c_void_p, _c_int, _c_int, _c_void_p, ctypes.c_longlong, _c_void_p, _c_int, _c_int, _c_int,
                 _c_int, _c_void_p]),
    "ebsd_sample_dirs_grad": (
        _c_int, [_c_void_p, _c_int, _c_int, _c_void_p, ctypes.c_longlong, _c_void_p, _c_void_p, _c_int, _c_int,
                 _c_int, _c_void_p]),
    "ebsd_prep_patterns": (
        _c_int,
        [_c_void_p, _c_int, _c_int, _c_void_p, _c_int, _c_void_p, _c_int, _c_int,
         _c_void_p, _c_int, _c_void_p],
    ),
    "ebsd_prep_patterns_pair": (
        _c_int,
        [_c_void_p, _c_int, _c_int, _c_void_p, _c_int, _c_void_p, _c_int, _c_void_p, _c_int, _c_int, _c_int,
         _c_void_p],
    ),
    "ebsd_rescore": (
        _c_int,
        [_c_void_p, _c_int, _c_void_p, _c_int, _c_int, _c_void_p, _c_int, _c_int, _c_void_p, _c_void_p, _c_void_p]),
    "ebsd_gemm_nt": (
        _c_int, [_c_void_p, _c_int, _c_void_p, _c_int, _c_int, _c_int, _c_void_p, _c_int, ctypes.c_float,
                 ctypes.c_float, _c_void_p]),
    "ebsd_gemm_nt_splitk_workspace": (_c_int, [_c_int, _c_int, _c_int, ctypes.POINTER(_c_size_t)]),
    "ebsd_gemm_nt_splitk": (
        _c_int, [_c_void_p, _c_int, _c_void_p, _c_int, _c_int, _c_int, _c_void_p, _c_int, ctypes.c_float,
                 ctypes.c_float, _c_int, _c_void_p, _c_size_t, _c_void_p]),
    "ebsd_cov_update_workspace": (_c_int, [_c_int, _c_int, ctypes.POINTER(_c_size_t)]),
    "ebsd_cov_update": (
        _c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_void_p, _c_void_p, _c_int, ctypes.c_longlong, _c_void_p,
                 _c_size_t, _c_void_p]),
    "ebsd_nn_im2col": (_c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_int, ctypes.c_float, _c_void_p, _c_int, _c_int,
                                _c_int, _c_void_p]),
    "ebsd_nn_rot6d_to_quat": (_c_int, [_c_void_p, _c_int, _c_int, _c_void_p, _c_void_p]),
    "ebsd_nn_cast": (_c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_void_p, _c_int, _c_int, _c_void_p]),
    "ebsd_mp_quad_table": (_c_int, [_c_void_p, _c_int, _c_int, _c_void_p, _c_int, _c_void_p]),
    "ebsd_local_refine": (
        _c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_void_p, _c_int, _c_void_p, _c_int, _c_int, _c_void_p,
                 _c_void_p, _c_int, _c_int, _c_void_p, _c_void_p, _c_void_p]),
    "ebsd_local_rescore": (
        _c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_void_p, _c_int, _c_void_p, _c_int, _c_int, _c_void_p,
                 _c_void_p, _c_int, _c_void_p, _c_int, _c_void_p, _c_void_p, _c_void_p, _c_void_p]),
    "ebsd_nn_conv_c1": (
        _c_int, [_c_void_p, _c_int, _c_int, _c_int, ctypes.c_float, _c_void_p, _c_int, _c_int, _c_void_p, _c_int,
                 _c_void_p, _c_void_p]),
    "ebsd_nn_conv_gemm": (
        _c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_int, _c_int, _c_int, _c_void_p,
                 _c_void_p, _c_void_p]),
    "ebsd_gemm_nt_bias_act": (
        _c_int, [_c_void_p, _c_int, _c_void_p, _c_int, _c_int, _c_int, _c_void_p, _c_int, _c_void_p, _c_int,
                 _c_void_p]),
    "ebsd_grid_refine_step": (
        _c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_void_p, _c_int, _c_void_p, _c_int, _c_int, _c_void_p,
                 _c_void_p, _c_int, _c_int, _c_void_p, _c_void_p, _c_int, _c_void_p, _c_void_p, _c_void_p,
                 _c_void_p]),
    "ebsd_nlpar_distances": (_c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_void_p, _c_void_p]),
    "ebsd_nlpar_average": (_c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_void_p, _c_void_p,
                                    _c_void_p]),
    "ebsd_clahe": (_c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, ctypes.c_float, _c_void_p,
                            _c_void_p]),
    "ebsd_normalize_patterns": (_c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_void_p]),
    "ebsd_nn_col2im": (_c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_void_p]),
    "ebsd_nn_bias_act": (_c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_void_p, _c_int, _c_void_p]),
    "ebsd_nn_relu_bwd": (_c_int, [_c_void_p, _c_void_p, ctypes.c_longlong, _c_void_p]),
    "ebsd_nn_colsum_workspace": (_c_int, [_c_int, _c_int, ctypes.POINTER(_c_size_t)]),
    "ebsd_nn_colsum": (_c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_void_p, _c_void_p, _c_size_t, _c_void_p]),
    "ebsd_nn_transpose": (_c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_void_p, _c_int, _c_int, _c_void_p]),
    "ebsd_nn_geodesic_loss": (
        _c_int, [_c_void_p, _c_int, _c_int, _c_void_p, _c_void_p, _c_int, _c_void_p, _c_void_p, _c_int, _c_void_p]),
    "ebsd_nn_adam": (
        _c_int, [_c_void_p, _c_void_p, _c_void_p, _c_void_p, ctypes.c_longlong, ctypes.c_float, ctypes.c_float,
                 ctypes.c_float, ctypes.c_float, _c_int, ctypes.c_float, _c_void_p, _c_void_p]),
    "ebsd_sim_topk_workspace": (_c_int, [_c_int, _c_int, _c_int, _c_int, ctypes.POINTER(_c_size_t)]),
    "ebsd_sim_schedule": (
        _c_int,
        [_c_int, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_void_p, ctypes.c_longlong, _c_void_p,
         ctypes.c_longlong],
    ),
    "ebsd_sim_topk": (
        _c_int,
        [_c_void_p, _c_void_p, _c_int, _c_void_p, _c_void_p, _c_int, _c_int, _c_int, _c_int,
         _c_void_p, _c_void_p, _c_int, _c_int, _c_void_p, _c_size_t, _c_void_p],
    ),
    "ebsd_sim_topk_noalloc": (
        _c_int,
        [_c_void_p, _c_void_p, _c_int, _c_void_p, _c_void_p, _c_int, _c_int, _c_int, _c_int,
         _c_void_p, _c_void_p, _c_int, _c_int, _c_void_p],
    ),
}

_lock = threading.Lock()
_lib = None


class EbsdLibraryError(RuntimeError):
    """The native library is missing, stale or returned a HIP error."""


def exported_symbols():
    return list(_SIGNATURES)


def load(path: str | None = None):
    """Load (once) and return the ctypes handle of libebsd_amd.so."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        so = path or os.environ.get("EBSD_AMD_LIB", LIB_PATH)
        if not os.path.exists(so):
            raise EbsdLibraryError(
                f"{so} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(hipcc, gfx950); this package has no non-HIP code path"
            )
        lib = ctypes.CDLL(so)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)  # AttributeError = stale library
            fn.restype = res
            fn.argtypes = args
        ver = lib.ebsd_abi_version()
        if ver != EBSD_ABI_VERSION:
            raise EbsdLibraryError(f"{so} has ABI {ver}, expected {EBSD_ABI_VERSION}: rebuild it")
        _lib = lib
        return lib


def check(status: int, what: str) -> None:
    """Map a status code to the exception types of the reference API."""
    if status == EBSD_OK:
        return
    lib = load()
    msg = f"{what}: {lib.ebsd_status_string(status).decode()} (status {status})"
    if status in (EBSD_ERR_SHAPE, EBSD_ERR_TOPK, EBSD_ERR_NULL, EBSD_ERR_ALIGN, EBSD_ERR_WORKSPACE,
                  EBSD_ERR_SCRATCH):
        raise ValueError(msg)
    if status == EBSD_ERR_DTYPE:
        raise NotImplementedError(msg)
    raise EbsdLibraryError(msg)


def stream_ptr(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def require_device(t: torch.Tensor, name: str) -> None:
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be a GPU tensor (got {t.device}); the HIP path has no CPU mode")
