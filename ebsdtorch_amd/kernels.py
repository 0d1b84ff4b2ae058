"""Thin Python wrappers over the C ABI (include/ebsd_amd.h).

Tensors cross the boundary as raw device pointers; every call is ordered on
torch's current HIP stream of the tensors' device. Shapes and dtypes are
checked here, before a pointer reaches a kernel, and the library re-checks
them and returns an error code instead of launching.
"""

from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import torch

from . import _lib
from ._lib import (
    DTYPE_BF16,
    DTYPE_F16,
    DTYPE_F32,
    DTYPE_FP8_E4M3,
    FLAG_RENORM_DIRS,
    FLAG_UNIT_NORM,
    FLAG_ZERO_MEAN,
    TORCH_DTYPE,
    check,
    require_device,
)

MATCH_DTYPES = {
    torch.float32: DTYPE_F32,
    torch.bfloat16: DTYPE_BF16,
    torch.float16: DTYPE_F16,
}
PRECISIONS = {"fp32": DTYPE_F32, "bf16": DTYPE_BF16, "fp16": DTYPE_F16, "fp8": DTYPE_FP8_E4M3}


def pad_multiple(dtype: int) -> int:
    return _lib.load().ebsd_pad_multiple(dtype)


def padded_width(n_pix: int, dtype: int) -> int:
    m = pad_multiple(dtype)
    return ((n_pix + m - 1) // m) * m


class PatternMatrix:
    """Rows of zero-padded patterns in one dtype, as the kernels consume them.

    data  (n, ld) tensor of TORCH_DTYPE[dtype] (uint8 bytes for FP8)
    scale (n,) float32 per-row scale (FP8: value = stored * scale), or None
    """

    __slots__ = ("data", "scale", "dtype", "n_pix")

    def __init__(self, data: torch.Tensor, scale: Optional[torch.Tensor], dtype: int, n_pix: int):
        self.data, self.scale, self.dtype, self.n_pix = data, scale, dtype, n_pix

    @property
    def rows(self) -> int:
        return self.data.shape[0]

    @property
    def ld(self) -> int:
        return self.data.shape[1]

    def narrow(self, start: int, length: int) -> "PatternMatrix":
        s = None if self.scale is None else self.scale.narrow(0, start, length)
        return PatternMatrix(self.data.narrow(0, start, length), s, self.dtype, self.n_pix)

    def to_float(self) -> torch.Tensor:
        """Dequantised float32 copy of the first n_pix columns."""
        d = self.data[:, : self.n_pix]
        if self.dtype == DTYPE_FP8_E4M3:
            f = d.view(torch.float8_e4m3fn).float()
        else:
            f = d.float()
        if self.scale is not None:
            f = f * self.scale[:, None]
        return f


def alloc_patterns(n: int, n_pix: int, dtype: int, device: torch.device) -> PatternMatrix:
    ld = padded_width(n_pix, dtype)
    data = torch.empty((n, ld), dtype=TORCH_DTYPE[dtype], device=device)
    scale = torch.empty((n,), dtype=torch.float32, device=device) if dtype == DTYPE_FP8_E4M3 else None
    return PatternMatrix(data, scale, dtype, n_pix)


def project(master_pattern: torch.Tensor, rays: torch.Tensor, quats: torch.Tensor, dtype: int = DTYPE_F32,
            zero_mean: bool = True, unit_norm: bool = False, renorm_dirs: bool = False,
            out: Optional[PatternMatrix] = None) -> PatternMatrix:
    """Dictionary projection (ebsd_project). master_pattern (S, 2S) f32,
    rays (D, 3) f32, quats (N, 4) f32 (w, x, y, z)."""
    for t, name in ((master_pattern, "master_pattern"), (rays, "rays"), (quats, "quats")):
        require_device(t, name)
        if t.dtype != torch.float32:
            raise ValueError(f"{name} must be float32")
    mp = master_pattern.contiguous()
    rays = rays.contiguous()
    quats = quats.contiguous()
    if mp.dim() != 2:
        raise ValueError(f"master pattern must be a 2D side-by-side image, got {tuple(mp.shape)}")
    if rays.dim() != 2 or rays.shape[1] != 3:
        raise ValueError("rays must be (D, 3)")
    if quats.dim() != 2 or quats.shape[1] != 4:
        raise ValueError("quats must be (N, 4)")
    n, d = quats.shape[0], rays.shape[0]
    if out is None:
        out = alloc_patterns(n, d, dtype, quats.device)
    elif out.rows != n or out.n_pix != d or out.dtype != dtype:
        raise ValueError("output buffer does not match the projection")
    flags = (FLAG_ZERO_MEAN if zero_mean else 0) | (FLAG_UNIT_NORM if unit_norm else 0) | (
        FLAG_RENORM_DIRS if renorm_dirs else 0)
    lib = _lib.load()
    st = lib.ebsd_project(mp.data_ptr(), mp.shape[0], mp.shape[1], rays.data_ptr(), d, quats.data_ptr(), n,
                          out.data.data_ptr(), dtype, out.ld, _lib.ptr(out.scale), flags,
                          _lib.stream_ptr(quats.device))
    check(st, "ebsd_project")
    return out


def prep_patterns(pats: torch.Tensor, dtype: int = DTYPE_F32, zero_mean: bool = True, unit_norm: bool = False,
                  pix_idx: Optional[torch.Tensor] = None, out: Optional[PatternMatrix] = None) -> PatternMatrix:
    """Experimental-pattern preparation (ebsd_prep_patterns).
    pats (N, P) float32 on the GPU; pix_idx (D,) int32 selects columns."""
    require_device(pats, "patterns")
    if pats.dtype != torch.float32:
        pats = pats.float()
    if pats.dim() != 2:
        raise ValueError("patterns must be (N, P)")
    if pats.stride(1) != 1:
        pats = pats.contiguous()
    n, p = pats.shape
    if pix_idx is not None:
        pix_idx = pix_idx.to(device=pats.device, dtype=torch.int32).contiguous()
        d = pix_idx.numel()
        if d == 0 or int(pix_idx.min()) < 0 or int(pix_idx.max()) >= p:
            raise ValueError("pixel indices must lie in [0, n_pixels)")
    else:
        d = p
    if out is None:
        out = alloc_patterns(n, d, dtype, pats.device)
    elif out.rows != n or out.n_pix != d or out.dtype != dtype:
        raise ValueError("output buffer does not match the patterns")
    flags = (FLAG_ZERO_MEAN if zero_mean else 0) | (FLAG_UNIT_NORM if unit_norm else 0)
    lib = _lib.load()
    st = lib.ebsd_prep_patterns(pats.data_ptr(), n, pats.stride(0), _lib.ptr(pix_idx), d, out.data.data_ptr(),
                                dtype, out.ld, _lib.ptr(out.scale), flags, _lib.stream_ptr(pats.device))
    check(st, "ebsd_prep_patterns")
    return out


_ws_cache: dict = {}


def _workspace(device: torch.device, nbytes: int) -> torch.Tensor:
    key = (device.type, device.index)
    buf = _ws_cache.get(key)
    if buf is None or buf.numel() < nbytes:
        buf = torch.empty(max(nbytes, 1 << 20), dtype=torch.uint8, device=device)
        _ws_cache[key] = buf
    return buf


def sim_topk(queries: PatternMatrix, dictionary: PatternMatrix, topk: int, top_val: torch.Tensor,
             top_idx: torch.Tensor, index_base: int = 0, merge: bool = True) -> None:
    """Streaming similarity top-k (ebsd_sim_topk): merges the best `topk`
    dictionary rows per query into (top_val, top_idx) in place."""
    if queries.dtype != dictionary.dtype:
        raise ValueError("queries and dictionary must share a dtype")
    if queries.ld != dictionary.ld:
        raise ValueError("queries and dictionary must share a leading dimension")
    m, n = queries.rows, dictionary.rows
    if top_val.shape != (m, topk) or top_idx.shape != (m, topk):
        raise ValueError("top-k state must be (n_queries, topk)")
    if top_val.dtype != torch.float32 or top_idx.dtype != torch.int32:
        raise ValueError("top-k state must be float32 values and int32 indices")
    if not (top_val.is_contiguous() and top_idx.is_contiguous()):
        raise ValueError("top-k state must be contiguous")
    lib = _lib.load()
    nbytes = ctypes.c_size_t(0)
    check(lib.ebsd_sim_topk_workspace(m, n, topk, queries.dtype, ctypes.byref(nbytes)), "ebsd_sim_topk_workspace")
    dev = queries.data.device
    ws = _workspace(dev, nbytes.value)
    st = lib.ebsd_sim_topk(queries.data.data_ptr(), _lib.ptr(queries.scale), m, dictionary.data.data_ptr(),
                           _lib.ptr(dictionary.scale), n, queries.ld, queries.dtype, topk, top_val.data_ptr(),
                           top_idx.data_ptr(), index_base, 1 if merge else 0, ws.data_ptr(), ws.numel(),
                           _lib.stream_ptr(dev))
    check(st, "ebsd_sim_topk")


def rescore(queries_f32: PatternMatrix, dictionary_f32: PatternMatrix, candidates: torch.Tensor, topk: int):
    """Exact FP32 re-scoring of candidate lists (ebsd_rescore): returns
    (values (N, topk) f32, indices (N, topk) int32), best first."""
    if queries_f32.dtype != DTYPE_F32 or dictionary_f32.dtype != DTYPE_F32:
        raise ValueError("rescoring runs on FP32 pattern matrices")
    if queries_f32.ld != dictionary_f32.ld:
        raise ValueError("queries and dictionary must share a leading dimension")
    cand = candidates.to(torch.int32).contiguous()
    m, c = cand.shape
    if m != queries_f32.rows:
        raise ValueError("one candidate row per query")
    dev = cand.device
    val = torch.empty((m, topk), dtype=torch.float32, device=dev)
    idx = torch.empty((m, topk), dtype=torch.int32, device=dev)
    st = _lib.load().ebsd_rescore(queries_f32.data.data_ptr(), m, dictionary_f32.data.data_ptr(),
                                  dictionary_f32.rows, queries_f32.ld,
                                  cand.data_ptr(), c, topk, val.data_ptr(), idx.data_ptr(), _lib.stream_ptr(dev))
    check(st, "ebsd_rescore")
    return val, idx


def gemm_nt(a: PatternMatrix, b: PatternMatrix, out: Optional[torch.Tensor] = None, alpha: float = 1.0,
            beta: float = 0.0) -> torch.Tensor:
    """C = alpha * A B^T + beta * C on the matrix cores (ebsd_gemm_nt):
    A (m, ld), B (n, ld) pattern matrices of one dtype -> C (m, n) float32
    (or written into `out`, whose row stride may exceed n)."""
    if a.dtype != b.dtype or a.ld != b.ld:
        raise ValueError("both operands need the same dtype and leading dimension")
    m, n = a.rows, b.rows
    dev = a.data.device
    if out is None:
        out = torch.empty((m, n), dtype=torch.float32, device=dev)
    if out.dtype != torch.float32 or out.shape[0] != m or out.shape[1] < n or out.stride(1) != 1:
        raise ValueError("output must be float32 (m, >= n) with unit column stride")
    st = _lib.load().ebsd_gemm_nt(a.data.data_ptr(), m, b.data.data_ptr(), n, a.ld, a.dtype, out.data_ptr(),
                                  out.stride(0), alpha, beta, _lib.stream_ptr(dev))
    check(st, "ebsd_gemm_nt")
    return out


def cov_update(x: PatternMatrix, mean: torch.Tensor, cov: torch.Tensor, n_seen: int) -> None:
    """Merge the rows of `x` (first n_pix columns, float32) into the running
    mean (d,) and covariance aggregate (d, d) in place (ebsd_cov_update);
    only block tiles on or above the diagonal of `cov` are maintained."""
    if x.dtype != DTYPE_F32:
        raise ValueError("covariance batches are float32")
    d = x.n_pix
    if mean.shape != (d,) or cov.shape[0] != d or cov.shape[1] < d:
        raise ValueError("mean / covariance shapes do not match the batch")
    if not (mean.is_contiguous() and cov.stride(1) == 1):
        raise ValueError("mean and covariance must be contiguous")
    lib = _lib.load()
    nbytes = ctypes.c_size_t(0)
    check(lib.ebsd_cov_update_workspace(x.rows, d, ctypes.byref(nbytes)), "ebsd_cov_update_workspace")
    dev = x.data.device
    ws = _workspace(dev, nbytes.value)
    st = lib.ebsd_cov_update(x.data.data_ptr(), x.rows, d, x.ld, mean.data_ptr(), cov.data_ptr(), cov.stride(0),
                             int(n_seen), ws.data_ptr(), ws.numel(), _lib.stream_ptr(dev))
    check(st, "ebsd_cov_update")
