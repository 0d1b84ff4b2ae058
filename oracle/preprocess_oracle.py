"""CPU oracle of ebsdtorch's per-pattern preprocessing: CLAHE
(preprocessing/clahe.py:18-130) and normalize_per_pattern
(ebsd/experiment_pats.py:256-281).

TEST INFRASTRUCTURE ONLY. Only tests/ may import this module, and only as
the checker; the product package never imports it.

CLAHE is restated tile by tile (reflect padding, torch.bincount histograms,
the clip limit and redistribution, the clamped CDF) and its final lookup is
torch's own 3D grid_sample, the library routine the reference calls. Pinned
in tests/test_oracle_preprocess.py (count conservation, CDF shape, global
equalisation as a monotone map, the lookup against a scalar trilinear loop).
"""

from __future__ import annotations

import math

import torch
import torch.nn.functional as F


def tile_cdfs(x: torch.Tensor, clip_limit: float = 40.0, n_bins: int = 64, grid_shape=(4, 4)) -> torch.Tensor:
    """(B, gh, gw, n_bins) clamped CDFs of the tiles of x (B, H, W)
    (clahe.py:41-102)."""
    b, h, w = x.shape
    gh, gw = grid_shape
    th, tw = math.ceil(h / gh), math.ceil(w / gw)
    vox = th * tw
    xp = F.pad(x[:, None], (0, gw * tw - w, 0, gh * th - h), mode="reflect")[:, 0]
    out = torch.empty((b, gh, gw, n_bins), dtype=torch.float32)
    limit = max(clip_limit * vox // n_bins, 1)
    for i in range(b):
        for ty in range(gh):
            for tx in range(gw):
                t = xp[i, ty * th:(ty + 1) * th, tx * tw:(tx + 1) * tw].reshape(-1)
                counts = torch.bincount((t * (n_bins - 1)).to(torch.int64), minlength=n_bins).to(torch.float32)
                hist = counts / counts.sum() * vox
                if clip_limit > 0:
                    hist = hist.clamp(max=limit)
                    clipped = vox - hist.sum()
                    residual = torch.remainder(clipped, n_bins)
                    hist = hist + (clipped - residual) / n_bins
                    hist = hist + (torch.arange(n_bins) < residual).to(torch.float32)
                out[i, ty, tx] = (torch.cumsum(hist, 0) / vox).clamp(0.0, 1.0)
    return out


def clahe_grayscale(x: torch.Tensor, clip_limit: float = 40.0, n_bins: int = 64, grid_shape=(4, 4)) -> torch.Tensor:
    """clahe.py:18-130 for images x (B, H, W) float32 in [0, 1]."""
    b, h, w = x.shape
    cdfs = tile_cdfs(x, clip_limit, n_bins, grid_shape)
    gx, gy = torch.meshgrid(torch.linspace(-1.0, 1.0, w), torch.linspace(-1.0, 1.0, h), indexing="xy")
    grid = torch.stack((2.0 * x - 1.0, gx.expand(b, h, w), gy.expand(b, h, w)), dim=-1)
    out = F.grid_sample(cdfs[:, None], grid[:, None], mode="bilinear", padding_mode="border", align_corners=False)
    return out[:, 0, 0]


def normalize_per_pattern(x: torch.Tensor, norm_type: str = "minmax") -> torch.Tensor:
    """experiment_pats.py:256-281 on (N, P) float32 (a new tensor)."""
    x = x.clone()
    if norm_type == "minmax":
        lo = x.min(dim=-1).values[:, None]
        hi = x.max(dim=-1).values[:, None]
        x -= lo
        x /= 1e-4 + hi - lo
    elif norm_type == "zeromean":
        x -= x.mean(dim=-1, keepdim=True)
    elif norm_type == "standard":
        x -= x.mean(dim=-1, keepdim=True)
        x /= x.std(dim=-1, keepdim=True)
    else:
        raise ValueError(norm_type)
    return x
