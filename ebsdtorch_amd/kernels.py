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
    scale (n,) float32 per-row scale (FP8: value = stored * scale, a power of
          two), or None
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


# projections sorted by view direction by default: at least this many rows
# and more than VIEW_ORDER_MIN_PIXELS pixels (measured on the 1M
# dictionaries: 120 x 120 70.7 -> 61.4 ms FP32, 86.0 -> 70.1 ms FP8; 80 x 80
# 30.8 -> 29.7 ms; at 60 x 60 the gain, 15.1 -> 14.8 ms, is less than the
# ~1.2 ms the sort of a 1M-entry key array costs)
VIEW_ORDER_MIN_ROWS = 1 << 16
VIEW_ORDER_MIN_PIXELS = 4096


def view_order(quats: torch.Tensor, rays: torch.Tensor) -> torch.Tensor:
    """Processing order of a projection (int32 permutation of the rows):
    orientations sorted by where they send the detector's central ray in the
    side-by-side Lambert image (ebsd_view_keys, Z-order cells), so the
    patterns projected together gather from the same master-pattern region."""
    quats = _f32_on_device(quats, "quats")
    c = rays.to(torch.float64).mean(0)
    c = (c / c.norm()).tolist()
    keys = torch.empty(quats.shape[0], dtype=torch.int32, device=quats.device)
    check(_lib.load().ebsd_view_keys(quats.data_ptr(), quats.shape[0], c[0], c[1], c[2], keys.data_ptr(),
                                     _lib.stream_ptr(quats.device)), "ebsd_view_keys")
    return torch.argsort(keys, stable=True).to(torch.int32)


def project(master_pattern: torch.Tensor, rays: torch.Tensor, quats: torch.Tensor, dtype: int = DTYPE_F32,
            zero_mean: bool = True, unit_norm: bool = False, renorm_dirs: bool = False,
            out: Optional[PatternMatrix] = None, quads: Optional[torch.Tensor] = None,
            order=None) -> PatternMatrix:
    """Dictionary projection (ebsd_project). master_pattern (S, 2S) f32,
    rays (D, 3) f32, quats (N, 4) f32 (w, x, y, z). With `quads`, the
    master pattern's float32 quad table (mp_quad_table), the gathers read it
    instead (ebsd_project_quads, same result bit for bit), in the processing
    `order` (an int32 permutation from view_order, or None; by default
    view_order's for N >= VIEW_ORDER_MIN_ROWS rows of more than
    VIEW_ORDER_MIN_PIXELS pixels): the output does not depend on it."""
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
    if quads is not None:
        require_device(quads, "quads")
        if quads.dtype != torch.float32 or quads.shape != (mp.shape[0], mp.shape[1], 4) or not quads.is_contiguous():
            raise ValueError("quads must be the float32 (S, C, 4) quad table of this master pattern")
        if order is None and n >= VIEW_ORDER_MIN_ROWS and d > VIEW_ORDER_MIN_PIXELS:
            order = view_order(quats, rays)
        if order is not None:
            require_device(order, "order")
            if order.dtype != torch.int32 or order.shape != (n,) or not order.is_contiguous():
                raise ValueError("order must be a contiguous (N,) int32 permutation")
        st = lib.ebsd_project_quads(quads.data_ptr(), mp.shape[0], mp.shape[1], rays.data_ptr(), d, quats.data_ptr(),
                                    n, out.data.data_ptr(), dtype, out.ld, _lib.ptr(out.scale), flags,
                                    _lib.ptr(order), _lib.stream_ptr(quats.device))
        check(st, "ebsd_project_quads")
        return out
    st = lib.ebsd_project(mp.data_ptr(), mp.shape[0], mp.shape[1], rays.data_ptr(), d, quats.data_ptr(), n,
                          out.data.data_ptr(), dtype, out.ld, _lib.ptr(out.scale), flags,
                          _lib.stream_ptr(quats.device))
    check(st, "ebsd_project")
    return out


def prep_patterns_pair(pats: torch.Tensor, dtype2: int, zero_mean: bool = True, unit_norm: bool = False,
                       pix_idx: Optional[torch.Tensor] = None):
    """prep_patterns to float32 and to dtype2 (BF16 / FP16) in one pass
    (ebsd_prep_patterns_pair): (float32 rows, dtype2 rows), the second equal
    bit for bit to prep_patterns(pats, dtype2, ...)."""
    require_device(pats, "patterns")
    if dtype2 not in (DTYPE_BF16, DTYPE_F16):
        raise ValueError("the second output is BF16 or FP16")
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
    out = alloc_patterns(n, d, DTYPE_F32, pats.device)
    out2 = alloc_patterns(n, d, dtype2, pats.device)
    flags = (FLAG_ZERO_MEAN if zero_mean else 0) | (FLAG_UNIT_NORM if unit_norm else 0)
    st = _lib.load().ebsd_prep_patterns_pair(pats.data_ptr(), n, pats.stride(0), _lib.ptr(pix_idx), d,
                                             out.data.data_ptr(), out.ld, out2.data.data_ptr(), dtype2, out2.ld,
                                             flags, _lib.stream_ptr(pats.device))
    check(st, "ebsd_prep_patterns_pair")
    return out, out2


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


_SCHEDULE_KEYS = ("bt", "grid", "n_parts", "k_list", "q_rows", "d_rows", "n_qtiles", "n_dtiles", "A", "B", "R")


def sim_schedule(m: int, n: int, topk: int, dtype: int, ld: int, counts: bool = True) -> dict:
    """The work plan of sim_topk for m queries x n dictionary rows of width
    ld (ebsd_sim_schedule), walked on the host with the kernels' own block ->
    work mapping; no GPU is needed. With counts, "tile_visits" (query tiles
    x dictionary tiles) and "list_writes" (query tiles x partial lists) are
    int32 arrays of how often each tile pair is scored and each partial list
    written."""
    import numpy as np

    lib = _lib.load()
    info = (ctypes.c_int * _lib.EBSD_SCHEDULE_INFO)()
    check(lib.ebsd_sim_schedule(m, n, topk, dtype, ld, info, None, 0, None, 0), "ebsd_sim_schedule")
    out = dict(zip(_SCHEDULE_KEYS, list(info)))
    if counts:
        visits = np.empty((out["n_qtiles"], out["n_dtiles"]), dtype=np.int32)
        writes = np.empty((out["n_qtiles"], out["n_parts"]), dtype=np.int32)
        check(lib.ebsd_sim_schedule(m, n, topk, dtype, ld, info, visits.ctypes.data, visits.size,
                                    writes.ctypes.data, writes.size), "ebsd_sim_schedule")
        out["tile_visits"], out["list_writes"] = visits, writes
    return out


def sim_topk_noalloc(queries: PatternMatrix, dictionary: PatternMatrix, topk: int, top_val: torch.Tensor,
                     top_idx: torch.Tensor, index_base: int = 0, merge: bool = True) -> None:
    """sim_topk through the allocation-free entry point
    (ebsd_sim_topk_noalloc): the partial lists use the library's static
    device buffer, no workspace is allocated. Sizes beyond that buffer raise
    ValueError; calls overlapping on one GPU must be serialised."""
    if queries.dtype != dictionary.dtype or queries.ld != dictionary.ld:
        raise ValueError("queries and dictionary must share a dtype and a leading dimension")
    m, n = queries.rows, dictionary.rows
    if top_val.shape != (m, topk) or top_idx.shape != (m, topk):
        raise ValueError("top-k state must be (n_queries, topk)")
    if top_val.dtype != torch.float32 or top_idx.dtype != torch.int32:
        raise ValueError("top-k state must be float32 values and int32 indices")
    if not (top_val.is_contiguous() and top_idx.is_contiguous()):
        raise ValueError("top-k state must be contiguous")
    lib = _lib.load()
    st = lib.ebsd_sim_topk_noalloc(queries.data.data_ptr(), _lib.ptr(queries.scale), m, dictionary.data.data_ptr(),
                                   _lib.ptr(dictionary.scale), n, queries.ld, queries.dtype, topk,
                                   top_val.data_ptr(), top_idx.data_ptr(), index_base, 1 if merge else 0,
                                   _lib.stream_ptr(queries.data.device))
    check(st, "ebsd_sim_topk_noalloc")


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


def _f32_on_device(t: torch.Tensor, name: str) -> torch.Tensor:
    require_device(t, name)
    if t.dtype != torch.float32:
        raise ValueError(f"{name} must be float32")
    return t.contiguous()


def project_shifted(master_pattern: torch.Tensor, coords: torch.Tensor, quats: torch.Tensor,
                    offsets: torch.Tensor, dtype: int = DTYPE_F32, zero_mean: bool = True,
                    unit_norm: bool = False) -> PatternMatrix:
    """ebsd_project with a per-pattern shift of the un-normalised pixel
    positions: ray d of pattern n is normalise(coords[d] + offsets[n])."""
    mp = _f32_on_device(master_pattern, "master_pattern")
    coords = _f32_on_device(coords, "coords")
    quats = _f32_on_device(quats, "quats")
    offsets = _f32_on_device(offsets, "offsets")
    if coords.dim() != 2 or coords.shape[1] != 3:
        raise ValueError("coords must be (D, 3)")
    n, d = quats.shape[0], coords.shape[0]
    if quats.shape != (n, 4) or offsets.shape != (n, 3):
        raise ValueError("quats must be (N, 4) and offsets (N, 3)")
    out = alloc_patterns(n, d, dtype, quats.device)
    flags = (FLAG_ZERO_MEAN if zero_mean else 0) | (FLAG_UNIT_NORM if unit_norm else 0)
    st = _lib.load().ebsd_project_shifted(mp.data_ptr(), mp.shape[0], mp.shape[1], coords.data_ptr(), d,
                                          quats.data_ptr(), offsets.data_ptr(), n, out.data.data_ptr(), dtype,
                                          out.ld, _lib.ptr(out.scale), flags, _lib.stream_ptr(quats.device))
    check(st, "ebsd_project_shifted")
    return out


def poisson_gaussian_noise(clean: torch.Tensor, n_pix: int, mean_counts: torch.Tensor, sigma: float, seed: int,
                           offset: int = 0, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Counter-based detector noise (ebsd_poisson_gaussian_noise) on the
    first n_pix columns of each row of `clean` (N, >= n_pix) f32."""
    require_device(clean, "clean")
    if clean.dtype != torch.float32 or clean.dim() != 2 or clean.stride(1) != 1:
        raise ValueError("clean patterns must be float32 (N, P) rows")
    n = clean.shape[0]
    mc = _f32_on_device(mean_counts, "mean_counts")
    if mc.shape != (n,):
        raise ValueError("mean_counts must be (N,)")
    if out is None:
        out = torch.empty((n, n_pix), dtype=torch.float32, device=clean.device)
    if out.dtype != torch.float32 or out.shape[0] != n or out.shape[1] < n_pix or out.stride(1) != 1:
        raise ValueError("output must be float32 (N, >= n_pix)")
    st = _lib.load().ebsd_poisson_gaussian_noise(clean.data_ptr(), n, n_pix, clean.stride(0), out.data_ptr(),
                                                 out.stride(0), mc.data_ptr(), float(sigma),
                                                 int(seed) & 0xFFFFFFFFFFFFFFFF, int(offset) & 0xFFFFFFFFFFFFFFFF,
                                                 _lib.stream_ptr(clean.device))
    check(st, "ebsd_poisson_gaussian_noise")
    return out


_DT_OF = {torch.float32: DTYPE_F32, torch.bfloat16: DTYPE_BF16, torch.float16: DTYPE_F16}


def matmul_nt(a: torch.Tensor, b: torch.Tensor, out: torch.Tensor, alpha: float = 1.0, beta: float = 0.0,
              splits: int = 1) -> torch.Tensor:
    """out (m, n) = alpha * a b^T + beta * out for row-major a (m, K),
    b (n, K) of one dtype, K a multiple of pad_multiple(dtype) with zero
    padding (ebsd_gemm_nt, or ebsd_gemm_nt_splitk when splits > 1)."""
    if a.dtype != b.dtype or a.dim() != 2 or b.dim() != 2 or a.shape[1] != b.shape[1]:
        raise ValueError("operands must be 2D with one dtype and one K")
    if not (a.is_contiguous() and b.is_contiguous()):
        raise ValueError("operands must be contiguous")
    dt = _DT_OF.get(a.dtype)
    if dt is None:
        raise ValueError(f"unsupported operand dtype {a.dtype}")
    m, n = a.shape[0], b.shape[0]
    if out.dtype != torch.float32 or out.shape[0] != m or out.shape[1] < n or out.stride(1) != 1:
        raise ValueError("output must be float32 (m, >= n) with unit column stride")
    lib = _lib.load()
    dev = a.device
    if splits > 1:
        nbytes = ctypes.c_size_t(0)
        check(lib.ebsd_gemm_nt_splitk_workspace(m, n, splits, ctypes.byref(nbytes)), "ebsd_gemm_nt_splitk_workspace")
        ws = _workspace(dev, nbytes.value)
        st = lib.ebsd_gemm_nt_splitk(a.data_ptr(), m, b.data_ptr(), n, a.shape[1], dt, out.data_ptr(),
                                     out.stride(0), alpha, beta, splits, ws.data_ptr(), ws.numel(),
                                     _lib.stream_ptr(dev))
        check(st, "ebsd_gemm_nt_splitk")
    else:
        st = lib.ebsd_gemm_nt(a.data_ptr(), m, b.data_ptr(), n, a.shape[1], dt, out.data_ptr(), out.stride(0),
                              alpha, beta, _lib.stream_ptr(dev))
        check(st, "ebsd_gemm_nt")
    return out


def im2col(x: torch.Tensor, out: torch.Tensor, scale: float = 1.0, transposed: bool = False) -> torch.Tensor:
    """3x3 / stride 2 / padding 1 patches of NHWC x (B, H, W, C) f32 into
    out: (B*Ho*Wo, ld) or, transposed, (9C, ld) of float32 / bfloat16."""
    x = _f32_on_device(x, "x")
    b, h, w, c = x.shape
    dt = _DT_OF.get(out.dtype)
    if dt not in (DTYPE_F32, DTYPE_BF16) or not out.is_contiguous() or out.dim() != 2:
        raise ValueError("im2col output must be a contiguous float32/bfloat16 matrix")
    rows = b * ((h + 1) // 2) * ((w + 1) // 2)
    need = (9 * c, rows) if transposed else (rows, 9 * c)
    if out.shape[0] != need[0] or out.shape[1] < need[1]:
        raise ValueError(f"im2col output must be ({need[0]}, >= {need[1]})")
    st = _lib.load().ebsd_nn_im2col(x.data_ptr(), b, h, w, c, float(scale), out.data_ptr(), out.shape[1], dt,
                                    1 if transposed else 0, _lib.stream_ptr(x.device))
    check(st, "ebsd_nn_im2col")
    return out


def col2im(dcol: torch.Tensor, shape: Tuple[int, int, int, int], out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Adjoint of im2col (ebsd_nn_col2im): dcol (B*Ho*Wo, ld) f32 -> (B, H, W, C)."""
    dcol = _f32_on_device(dcol, "dcol")
    b, h, w, c = shape
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=dcol.device)
    if dcol.shape[0] != b * ((h + 1) // 2) * ((w + 1) // 2) or dcol.shape[1] < 9 * c:
        raise ValueError("dcol does not match the image shape")
    st = _lib.load().ebsd_nn_col2im(dcol.data_ptr(), dcol.shape[1], b, h, w, c, out.data_ptr(),
                                    _lib.stream_ptr(dcol.device))
    check(st, "ebsd_nn_col2im")
    return out


def cast_pad(src: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
    """out (rows, ld) float32/bfloat16 = src (rows, cols) f32, zero columns beyond."""
    require_device(src, "src")
    if src.dtype != torch.float32 or src.dim() != 2 or src.stride(1) != 1:
        raise ValueError("src must be float32 rows")
    dt = _DT_OF.get(out.dtype)
    if dt not in (DTYPE_F32, DTYPE_BF16) or not out.is_contiguous() or out.shape[0] != src.shape[0] \
            or out.shape[1] < src.shape[1]:
        raise ValueError("cast output must be contiguous (rows, >= cols) float32/bfloat16")
    st = _lib.load().ebsd_nn_cast(src.data_ptr(), src.shape[0], src.shape[1], src.stride(0), out.data_ptr(),
                                  out.shape[1], dt, _lib.stream_ptr(src.device))
    check(st, "ebsd_nn_cast")
    return out


def transpose_pad(src: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
    """out (cols, ld) float32/bfloat16 = src (rows, cols)^T, zero columns
    rows..ld-1."""
    require_device(src, "src")
    if src.dtype != torch.float32 or src.dim() != 2 or src.stride(1) != 1:
        raise ValueError("src must be float32 rows")
    dt = _DT_OF.get(out.dtype)
    rows, cols = src.shape
    if dt not in (DTYPE_F32, DTYPE_BF16) or not out.is_contiguous() or out.shape[0] != cols or out.shape[1] < rows:
        raise ValueError("transpose output must be contiguous (cols, >= rows)")
    st = _lib.load().ebsd_nn_transpose(src.data_ptr(), rows, cols, src.stride(0), out.data_ptr(), out.shape[1], dt,
                                       _lib.stream_ptr(src.device))
    check(st, "ebsd_nn_transpose")
    return out


def matmul_nt_bias_act(a: torch.Tensor, b: torch.Tensor, out: torch.Tensor, bias: torch.Tensor,
                       relu: bool) -> torch.Tensor:
    """out (m, n) = act(a b^T + bias) in one pass (ebsd_gemm_nt_bias_act);
    operands as matmul_nt, act = ReLU if relu."""
    if a.dtype != b.dtype or a.dim() != 2 or b.dim() != 2 or a.shape[1] != b.shape[1]:
        raise ValueError("operands must be 2D with one dtype and one K")
    if not (a.is_contiguous() and b.is_contiguous()):
        raise ValueError("operands must be contiguous")
    dt = _DT_OF.get(a.dtype)
    if dt is None:
        raise ValueError(f"unsupported operand dtype {a.dtype}")
    m, n = a.shape[0], b.shape[0]
    if out.dtype != torch.float32 or out.shape[0] != m or out.shape[1] < n or out.stride(1) != 1:
        raise ValueError("output must be float32 (m, >= n) with unit column stride")
    bias = _f32_on_device(bias, "bias")
    if bias.numel() < n:
        raise ValueError("bias needs one value per output column")
    st = _lib.load().ebsd_gemm_nt_bias_act(a.data_ptr(), m, b.data_ptr(), n, a.shape[1], dt, out.data_ptr(),
                                           out.stride(0), bias.data_ptr(), 1 if relu else 0,
                                           _lib.stream_ptr(a.device))
    check(st, "ebsd_gemm_nt_bias_act")
    return out


def conv_c1(x: torch.Tensor, weight: torch.Tensor, bias: torch.Tensor, scale: float = 1.0,
            out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """First convolution (one input channel, 3x3 / stride 2 / padding 1) with
    bias + ReLU, done directly (ebsd_nn_conv_c1): x (B, H, W, 1) or (B, H, W)
    f32, weight (C_out, >= 9) f32/bf16 with columns ky*3 + kx -> (B*Ho*Wo,
    C_out) f32."""
    x = _f32_on_device(x, "x")
    if x.dim() not in (3, 4) or (x.dim() == 4 and x.shape[3] != 1):
        raise ValueError("conv_c1 takes one-channel images (B, H, W[, 1])")
    b, h, w = x.shape[0], x.shape[1], x.shape[2]
    dt = _DT_OF.get(weight.dtype)
    if dt not in (DTYPE_F32, DTYPE_BF16) or weight.dim() != 2 or weight.stride(1) != 1 or weight.shape[1] < 9:
        raise ValueError("weight must be (C_out, >= 9) float32/bfloat16 with unit column stride")
    require_device(weight, "weight")
    c_out = weight.shape[0]
    bias = _f32_on_device(bias, "bias")
    rows = b * ((h + 1) // 2) * ((w + 1) // 2)
    if out is None:
        out = torch.empty((rows, c_out), dtype=torch.float32, device=x.device)
    if out.shape != (rows, c_out) or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError("out must be contiguous float32 (B*Ho*Wo, C_out)")
    st = _lib.load().ebsd_nn_conv_c1(x.data_ptr(), b, h, w, float(scale), weight.data_ptr(), weight.stride(0), dt,
                                     bias.data_ptr(), c_out, out.data_ptr(), _lib.stream_ptr(x.device))
    check(st, "ebsd_nn_conv_c1")
    return out


def conv_gemm(x: torch.Tensor, weight: torch.Tensor, bias: torch.Tensor,
              out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """3x3 / stride 2 / padding 1 convolution + bias + ReLU as an implicit GEMM
    (ebsd_nn_conv_gemm): x (B, H, W, C) f32 NHWC, weight (C_out, K_pad)
    f32/bf16 with column (ky*3 + kx)*C + c, zero beyond 9C -> (B*Ho*Wo, C_out)
    f32, equal to im2col + matmul_nt_bias_act."""
    x = _f32_on_device(x, "x")
    if x.dim() != 4:
        raise ValueError("x must be NHWC (B, H, W, C)")
    b, h, w, c = x.shape
    dt = _DT_OF.get(weight.dtype)
    if dt not in (DTYPE_F32, DTYPE_BF16) or weight.dim() != 2 or not weight.is_contiguous():
        raise ValueError("weight must be a contiguous float32/bfloat16 (C_out, K_pad) matrix")
    require_device(weight, "weight")
    c_out = weight.shape[0]
    bias = _f32_on_device(bias, "bias")
    if bias.numel() < c_out:
        raise ValueError("bias needs one value per output channel")
    rows = b * ((h + 1) // 2) * ((w + 1) // 2)
    if out is None:
        out = torch.empty((rows, c_out), dtype=torch.float32, device=x.device)
    if out.shape != (rows, c_out) or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError("out must be contiguous float32 (B*Ho*Wo, C_out)")
    st = _lib.load().ebsd_nn_conv_gemm(x.data_ptr(), b, h, w, c, weight.data_ptr(), weight.shape[1], c_out, dt,
                                       bias.data_ptr(), out.data_ptr(), _lib.stream_ptr(x.device))
    check(st, "ebsd_nn_conv_gemm")
    return out


def bias_act(y: torch.Tensor, bias: torch.Tensor, cols: int, relu: bool) -> torch.Tensor:
    """y[:, :cols] = act(y[:, :cols] + bias) in place."""
    require_device(y, "y")
    bias = _f32_on_device(bias, "bias")
    if y.dtype != torch.float32 or y.dim() != 2 or y.stride(1) != 1 or bias.numel() < cols:
        raise ValueError("bias_act needs float32 rows and a bias per column")
    st = _lib.load().ebsd_nn_bias_act(y.data_ptr(), y.shape[0], cols, y.stride(0), bias.data_ptr(),
                                      1 if relu else 0, _lib.stream_ptr(y.device))
    check(st, "ebsd_nn_bias_act")
    return y


def relu_bwd(dy: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """dy *= (y > 0) in place (contiguous tensors of one shape)."""
    if dy.shape != y.shape or not (dy.is_contiguous() and y.is_contiguous()) or dy.dtype != torch.float32:
        raise ValueError("relu_bwd needs contiguous float32 tensors of one shape")
    st = _lib.load().ebsd_nn_relu_bwd(dy.data_ptr(), y.data_ptr(), dy.numel(), _lib.stream_ptr(dy.device))
    check(st, "ebsd_nn_relu_bwd")
    return dy


def colsum(x: torch.Tensor, cols: int, out: torch.Tensor) -> torch.Tensor:
    """out[c] = sum over rows of x[:, c] for c < cols (deterministic order)."""
    require_device(x, "x")
    if x.dtype != torch.float32 or x.dim() != 2 or x.stride(1) != 1 or out.numel() < cols:
        raise ValueError("colsum needs float32 rows")
    lib = _lib.load()
    nbytes = ctypes.c_size_t(0)
    check(lib.ebsd_nn_colsum_workspace(x.shape[0], cols, ctypes.byref(nbytes)), "ebsd_nn_colsum_workspace")
    ws = _workspace(x.device, nbytes.value)
    st = lib.ebsd_nn_colsum(x.data_ptr(), x.shape[0], cols, x.stride(0), out.data_ptr(), ws.data_ptr(), ws.numel(),
                            _lib.stream_ptr(x.device))
    check(st, "ebsd_nn_colsum")
    return out


def geodesic_loss(z6: torch.Tensor, q_true: torch.Tensor, ops: torch.Tensor, loss: torch.Tensor,
                  grad: torch.Tensor) -> None:
    """Symmetric geodesic loss of 6D outputs z6 (B, >= 6) against unit
    quaternions q_true (B, 4) over symmetry rotations ops (n, 4): per-sample
    loss (B,) and d(mean loss)/d(z6) into grad (B, >= 6)."""
    require_device(z6, "z6")
    q_true = _f32_on_device(q_true, "q_true")
    ops = _f32_on_device(ops, "ops")
    b = z6.shape[0]
    if z6.dtype != torch.float32 or z6.stride(1) != 1 or z6.shape[1] < 6 or q_true.shape != (b, 4):
        raise ValueError("z6 must be float32 (B, >= 6) and q_true (B, 4)")
    if loss.shape != (b,) or grad.shape[0] != b or grad.shape[1] < 6 or grad.stride(1) != 1:
        raise ValueError("loss must be (B,) and grad (B, >= 6)")
    st = _lib.load().ebsd_nn_geodesic_loss(z6.data_ptr(), z6.stride(0), b, q_true.data_ptr(), ops.data_ptr(),
                                           ops.shape[0], loss.data_ptr(), grad.data_ptr(), grad.stride(0),
                                           _lib.stream_ptr(z6.device))
    check(st, "ebsd_nn_geodesic_loss")


def adam_step(w: torch.Tensor, g: torch.Tensor, m: torch.Tensor, v: torch.Tensor, lr: float, beta1: float,
              beta2: float, eps: float, step: int, weight_decay: float = 0.0,
              w_bf16: Optional[torch.Tensor] = None) -> None:
    """One Adam update of the float32 master weights w (and their bfloat16
    copy) from gradient g with moment buffers m, v (all contiguous)."""
    for t in (w, g, m, v):
        if t.dtype != torch.float32 or not t.is_contiguous() or t.numel() != w.numel():
            raise ValueError("Adam buffers must be contiguous float32 of one size")
    if w_bf16 is not None and (w_bf16.dtype != torch.bfloat16 or w_bf16.numel() != w.numel()
                               or not w_bf16.is_contiguous()):
        raise ValueError("the bfloat16 copy must match the weights")
    st = _lib.load().ebsd_nn_adam(w.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), w.numel(), lr, beta1, beta2,
                                  eps, int(step), weight_decay, _lib.ptr(w_bf16), _lib.stream_ptr(w.device))
    check(st, "ebsd_nn_adam")


def mp_quad_table(master_pattern: torch.Tensor, dtype: torch.dtype = torch.float32) -> torch.Tensor:
    """(S, C) master pattern -> (S, C, 4) bilinear quad table in float32 or
    float16 (ebsd_mp_quad_table)."""
    mp = _f32_on_device(master_pattern, "master_pattern")
    if mp.dim() != 2:
        raise ValueError("master pattern must be a 2D side-by-side image")
    dt = {torch.float32: DTYPE_F32, torch.float16: DTYPE_F16}.get(dtype)
    if dt is None:
        raise ValueError("the quad table is float32 or float16")
    out = torch.empty((mp.shape[0], mp.shape[1], 4), dtype=dtype, device=mp.device)
    st = _lib.load().ebsd_mp_quad_table(mp.data_ptr(), mp.shape[0], mp.shape[1], out.data_ptr(), dt,
                                        _lib.stream_ptr(mp.device))
    check(st, "ebsd_mp_quad_table")
    return out


def local_refine(quads: torch.Tensor, rays: torch.Tensor, queries: PatternMatrix, centers: torch.Tensor,
                 grid: torch.Tensor, topk: int = 1) -> Tuple[torch.Tensor, torch.Tensor]:
    """Score the local orientations qu_prod(grid[j], centers[i]) of every
    query against it on the fly (ebsd_local_refine). quads: mp_quad_table of
    the master pattern (float32 or float16); queries: FP32 unit-norm pattern
    rows. Returns (cosine (N, topk) f32, grid index (N, topk) int32)."""
    if queries.dtype != DTYPE_F32:
        raise ValueError("local refinement scores FP32 queries")
    require_device(quads, "quads")
    qdt = {torch.float32: DTYPE_F32, torch.float16: DTYPE_F16}.get(quads.dtype)
    if qdt is None or quads.dim() != 3 or quads.shape[2] != 4 or not quads.is_contiguous():
        raise ValueError("quads must be the (S, C, 4) table of mp_quad_table")
    rays = _f32_on_device(rays, "rays")
    centers = _f32_on_device(centers, "centers")
    grid = _f32_on_device(grid, "grid")
    n = queries.rows
    if rays.shape != (queries.n_pix, 3) or centers.shape != (n, 4) or grid.dim() != 2 or grid.shape[1] != 4:
        raise ValueError("rays (D, 3), centers (N, 4) and grid (G, 4) must match the queries")
    dev = centers.device
    val = torch.empty((n, topk), dtype=torch.float32, device=dev)
    idx = torch.empty((n, topk), dtype=torch.int32, device=dev)
    st = _lib.load().ebsd_local_refine(quads.data_ptr(), qdt, quads.shape[0], quads.shape[1], rays.data_ptr(),
                                       queries.n_pix, queries.data.data_ptr(), n, queries.ld, centers.data_ptr(),
                                       grid.data_ptr(), grid.shape[0], topk, val.data_ptr(), idx.data_ptr(),
                                       _lib.stream_ptr(dev))
    check(st, "ebsd_local_refine")
    return val, idx


def local_rescore(quads: torch.Tensor, rays: torch.Tensor, queries: PatternMatrix, centers: torch.Tensor,
                  grid: torch.Tensor, cand: torch.Tensor, all_scores: bool = False):
    """Rescore listed candidates qu_prod(grid[cand[i, c]], centers[i]) of every
    query at the queries' own resolution (ebsd_local_rescore). cand: (N, k)
    grid indices, k <= 64, -1 = none. Returns (best cosine (N,) f32, best grid
    index (N,) int32, every score (N, k) f32 or None)."""
    if queries.dtype != DTYPE_F32:
        raise ValueError("local rescoring scores FP32 queries")
    require_device(quads, "quads")
    qdt = {torch.float32: DTYPE_F32, torch.float16: DTYPE_F16}.get(quads.dtype)
    if qdt is None or quads.dim() != 3 or quads.shape[2] != 4 or not quads.is_contiguous():
        raise ValueError("quads must be the (S, C, 4) table of mp_quad_table")
    rays = _f32_on_device(rays, "rays")
    centers = _f32_on_device(centers, "centers")
    grid = _f32_on_device(grid, "grid")
    n = queries.rows
    if rays.shape != (queries.n_pix, 3) or centers.shape != (n, 4) or grid.dim() != 2 or grid.shape[1] != 4:
        raise ValueError("rays (D, 3), centers (N, 4) and grid (G, 4) must match the queries")
    if cand.dim() != 2 or cand.shape[0] != n or not 1 <= cand.shape[1] <= 64:
        raise ValueError("cand must be (N, k) with 1 <= k <= 64")
    require_device(cand, "cand")
    cand = cand.to(torch.int32).contiguous()
    dev = centers.device
    k = cand.shape[1]
    best_val = torch.empty(n, dtype=torch.float32, device=dev)
    best_idx = torch.empty(n, dtype=torch.int32, device=dev)
    scores = torch.empty((n, k), dtype=torch.float32, device=dev) if all_scores else None
    st = _lib.load().ebsd_local_rescore(quads.data_ptr(), qdt, quads.shape[0], quads.shape[1], rays.data_ptr(),
                                        queries.n_pix, queries.data.data_ptr(), n, queries.ld, centers.data_ptr(),
                                        grid.data_ptr(), grid.shape[0], cand.data_ptr(), k,
                                        scores.data_ptr() if scores is not None else None, best_val.data_ptr(),
                                        best_idx.data_ptr(), _lib.stream_ptr(dev))
    check(st, "ebsd_local_rescore")
    return best_val, best_idx, scores


def grid_refine_step(quads: torch.Tensor, rays: torch.Tensor, queries: PatternMatrix, centers: torch.Tensor,
                     grids: torch.Tensor, level: torch.Tensor, ops: torch.Tensor):
    """One grid-refinement iteration (ebsd_grid_refine_step). quads: quad
    table of the (binned) master pattern; queries: zero-mean FP32 rows;
    grids (L, G, 4) local rotations per shrink level; level (N,) int32;
    ops (n_ops, 4) symmetry operators. Returns (best dot (N,) f32, grid index
    (N,) int32, fundamental-zone quaternion (N, 4) f32)."""
    if queries.dtype != DTYPE_F32:
        raise ValueError("grid refinement scores FP32 queries")
    require_device(quads, "quads")
    qdt = {torch.float32: DTYPE_F32, torch.float16: DTYPE_F16}.get(quads.dtype)
    if qdt is None or quads.dim() != 3 or quads.shape[2] != 4 or not quads.is_contiguous():
        raise ValueError("quads must be the (S, C, 4) table of mp_quad_table")
    rays = _f32_on_device(rays, "rays")
    centers = _f32_on_device(centers, "centers")
    grids = _f32_on_device(grids, "grids")
    ops = _f32_on_device(ops, "ops")
    n = queries.rows
    if rays.shape != (queries.n_pix, 3) or centers.shape != (n, 4):
        raise ValueError("rays (D, 3) and centers (N, 4) must match the queries")
    if grids.dim() != 3 or grids.shape[2] != 4 or ops.dim() != 2 or ops.shape[1] != 4:
        raise ValueError("grids must be (L, G, 4) and ops (n_ops, 4)")
    require_device(level, "level")
    level = level.to(torch.int32).contiguous()
    if level.shape != (n,):
        raise ValueError("level must be (N,)")
    dev = centers.device
    best_val = torch.empty(n, dtype=torch.float32, device=dev)
    best_idx = torch.empty(n, dtype=torch.int32, device=dev)
    best_q = torch.empty((n, 4), dtype=torch.float32, device=dev)
    st = _lib.load().ebsd_grid_refine_step(quads.data_ptr(), qdt, quads.shape[0], quads.shape[1], rays.data_ptr(),
                                           queries.n_pix, queries.data.data_ptr(), n, queries.ld, centers.data_ptr(),
                                           grids.data_ptr(), grids.shape[1], grids.shape[0], level.data_ptr(),
                                           ops.data_ptr(), ops.shape[0], best_val.data_ptr(), best_idx.data_ptr(),
                                           best_q.data_ptr(), _lib.stream_ptr(dev))
    check(st, "ebsd_grid_refine_step")
    return best_val, best_idx, best_q


def rot6d_to_quat(z6: torch.Tensor) -> torch.Tensor:
    """(B, >= 6) network outputs -> (B, 4) unit quaternions, w >= 0, of the
    Gram-Schmidt rotation the geodesic loss scores."""
    require_device(z6, "z6")
    if z6.dtype != torch.float32 or z6.dim() != 2 or z6.shape[1] < 6 or z6.stride(1) != 1:
        raise ValueError("z6 must be float32 (B, >= 6)")
    q = torch.empty((z6.shape[0], 4), dtype=torch.float32, device=z6.device)
    st = _lib.load().ebsd_nn_rot6d_to_quat(z6.data_ptr(), z6.stride(0), z6.shape[0], q.data_ptr(),
                                           _lib.stream_ptr(z6.device))
    check(st, "ebsd_nn_rot6d_to_quat")
    return q


def cmif_maps(corr: torch.Tensor) -> torch.Tensor:
    """Mutual information per shift (ebsd_cmif_maps) from the 5 x 5
    level-indicator cross-correlations corr (..., 5, 5, H, W) float32 on the
    GPU: returns (..., H, W)."""
    require_device(corr, "corr")
    if corr.dtype != torch.float32 or corr.dim() < 4 or corr.shape[-4:-2] != (5, 5):
        raise ValueError("corr must be float32 (..., 5, 5, H, W)")
    corr = corr.contiguous()
    lead = corr.shape[:-4]
    h, w = corr.shape[-2:]
    n_maps = int(torch.tensor(lead).prod().item()) if len(lead) else 1
    out = torch.empty(lead + (h, w), dtype=torch.float32, device=corr.device)
    check(_lib.load().ebsd_cmif_maps(corr.data_ptr(), n_maps, h * w, out.data_ptr(), _lib.stream_ptr(corr.device)),
          "ebsd_cmif_maps")
    return out
