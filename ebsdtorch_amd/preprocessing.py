"""Pattern preprocessing: NLPAR (preprocessing/nlpar.py), CLAHE
(preprocessing/clahe.py) and per-pattern normalisation
(ebsd/experiment_pats.py:256-281).

`nlpar` mirrors nlpar.py:20-192 (Brewick, Wright & Rowenhorst 2019): the
squared distance of every pattern to each pattern of its (2k+1)^2 window
(ebsd_nlpar_distances), a per-point noise estimate from the 8 nearest, the
normalised distances turned into logits whose temperature is fitted by
Newton's method so the mean softmax weight of the centre over the 3x3 window
equals `coeff`, and the softmax-weighted average of the window
(ebsd_nlpar_average). The two passes over the patterns are the HIP kernels;
the weights in between, (points x window) values, are the reference's own
expressions on the device.
"""

from __future__ import annotations

import torch

from . import _lib
from ._lib import check, require_device


def nlpar_offsets(k_rad: int) -> torch.Tensor:
    """(n, 2) int64 window offsets (dy, dx) in the reference's order:
    meshgrid 'ij' sorted by L1 length with torch.argsort on the CPU
    (nlpar.py:83-98); index 0 is (0, 0) and 1..8 the nearest."""
    offs = torch.stack(torch.meshgrid(torch.arange(-k_rad, k_rad + 1), torch.arange(-k_rad, k_rad + 1),
                                      indexing="ij"), dim=-1).reshape(-1, 2)
    return offs[torch.argsort(torch.sum(torch.abs(offs), dim=-1))]


def _scan(data: torch.Tensor, k_rad: int) -> torch.Tensor:
    require_device(data, "data")
    if data.dim() != 4:
        raise ValueError("The EBSD pattern data tensor must have 4 dimensions.")
    if not 1 <= k_rad <= 5:
        raise ValueError("k_rad must be in [1, 5]")
    return data.to(torch.float32).contiguous()


def nlpar_distances(data: torch.Tensor, k_rad: int) -> torch.Tensor:
    """(H_s, W_s, n_off) float32: squared distance of each pattern of the
    scan data (H_s, W_s, h, w) to the patterns at the window offsets
    (nlpar_offsets order; zero patterns outside the scan), nlpar.py:105-117."""
    x = _scan(data, k_rad)
    h_scan, w_scan, h_pat, w_pat = x.shape
    offs = nlpar_offsets(k_rad).to(x.device, torch.int32).contiguous()
    dists = torch.empty((h_scan, w_scan, offs.shape[0]), dtype=torch.float32, device=x.device)
    check(_lib.load().ebsd_nlpar_distances(x.data_ptr(), h_scan, w_scan, h_pat * w_pat, k_rad, offs.data_ptr(),
                                           dists.data_ptr(), _lib.stream_ptr(x.device)), "ebsd_nlpar_distances")
    return dists


def nlpar(data: torch.Tensor, k_rad: int = 5, coeff: float = 0.375, coeff_fit_tol: float = 1e-6,
          center_logit_bias: float = 0.0) -> torch.Tensor:
    """NLPAR denoising of a scan data (H_s, W_s, h, w) float32 on the GPU
    (nlpar.py:20-192); k_rad in [1, 5]. Returns the denoised scan."""
    x = _scan(data, k_rad)
    h_scan, w_scan, h_pat, w_pat = x.shape
    n_pix = h_pat * w_pat
    dev = x.device
    offs = nlpar_offsets(k_rad)
    n_off = offs.shape[0]
    offs_dev = offs.to(dev, torch.int32).contiguous()
    dists = nlpar_distances(x, k_rad)

    # the reference's padded-point mask (nlpar.py:50-59, :109-117)
    iy = torch.arange(h_scan, device=dev)[:, None, None] + offs_dev[:, 0].long()
    ix = torch.arange(w_scan, device=dev)[None, :, None] + offs_dev[:, 1].long()
    out_of_bounds = (iy < 0) | (iy >= h_scan) | (ix < 0) | (ix >= w_scan)

    # noise level per point and normalised distances (nlpar.py:119-138)
    # (every offset at once: the same per-element expression as the
    # reference's loop over the window)
    sigma_hat = torch.sqrt(torch.min(dists[:, :, 1:9], dim=-1).values / (2 * n_pix))
    sigma_hat_padded = torch.nn.functional.pad(sigma_hat, (k_rad,) * 4, mode="constant", value=1.0)
    sigma_other = sigma_hat_padded[iy + k_rad, ix + k_rad]
    s2 = sigma_hat[:, :, None] ** 2 + sigma_other ** 2
    dists = (dists - n_pix * s2) / (2 ** 0.5 * n_pix ** 0.5 * s2)
    weights = -dists.clamp_min_(0.0)

    # temperature fitted by Newton's method (nlpar.py:143-171)
    # (the temperature becomes a float32 device scalar after the first step,
    # as in the reference)
    guess = 1.0
    nn_weights = weights[:, :, :9]
    for _ in range(10):
        gsq = guess ** 2
        obj = torch.mean(torch.softmax(nn_weights / gsq, dim=-1)[:, :, 0]) - coeff
        e = torch.exp(nn_weights / gsq)
        grad = torch.mean(
            -2 * torch.exp(nn_weights[:, :, 0] / gsq)
            * (torch.sum(nn_weights[:, :, [0]] * e, dim=-1) - torch.sum(nn_weights * e, dim=-1))
            / (guess ** 3 * torch.sum(e, dim=-1) ** 2))
        guess = guess - obj / grad
        if bool(torch.abs(obj) < coeff_fit_tol):
            break
    weights /= guess ** 2
    if center_logit_bias > 0:
        weights[:, :, 0] += center_logit_bias
    weights = weights.masked_fill_(out_of_bounds, -torch.inf)
    weights = torch.nn.functional.softmax(weights, dim=-1).contiguous()

    out = torch.empty_like(x)
    check(_lib.load().ebsd_nlpar_average(x.data_ptr(), h_scan, w_scan, n_pix, k_rad, offs_dev.data_ptr(),
                                         weights.data_ptr(), out.data_ptr(), _lib.stream_ptr(dev)),
          "ebsd_nlpar_average")
    return out


NORM_MODES = {"minmax": 0, "zeromean": 1, "standard": 2}


def normalize_patterns(x: torch.Tensor, norm_type: str = "minmax") -> torch.Tensor:
    """Normalise each pattern (row of x viewed as (N, -1), float32 on the GPU)
    in place: "minmax" (x - min) / (1e-4 + max - min), "zeromean", or
    "standard" (zero mean, unit unbiased std) (experiment_pats.py:256-281)."""
    require_device(x, "x")
    if norm_type not in NORM_MODES:
        raise ValueError(f"Invalid normalization method. Got {norm_type} but expected 'minmax', "
                         "'zeromean', or 'standard'.")
    if x.dtype != torch.float32 or not x.is_contiguous():
        raise ValueError("patterns must be contiguous float32")
    n = x.shape[0]
    check(_lib.load().ebsd_normalize_patterns(x.data_ptr(), n, x.numel() // max(n, 1), NORM_MODES[norm_type],
                                              _lib.stream_ptr(x.device)), "ebsd_normalize_patterns")
    return x


def clahe_grayscale(x: torch.Tensor, clip_limit: float = 40.0, n_bins: int = 64, grid_shape=(4, 4)) -> torch.Tensor:
    """CLAHE of images x (..., H, W) float32 on the GPU, values in [0, 1]
    (clahe.py:18-130). Returns a new tensor of the same shape."""
    require_device(x, "x")
    h, w = x.shape[-2:]
    src = x.to(torch.float32).contiguous()
    n = src.numel() // (h * w)
    out = torch.empty_like(src)
    check(_lib.load().ebsd_clahe(src.data_ptr(), n, h, w, int(grid_shape[0]), int(grid_shape[1]), int(n_bins),
                                 float(clip_limit), out.data_ptr(), _lib.stream_ptr(src.device)), "ebsd_clahe")
    return out
