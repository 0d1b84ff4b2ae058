"""Seeded synthetic inputs for the DI path: a Laue-symmetric master pattern
built from Gaussian Kikuchi-like bands, and noisy experimental patterns
simulated from it with the HIP projection.

Master pattern (S, 2S): pixel (r, c) of the northern half sits at Lambert
square coordinates X = (2c + 1)/S - 1, Y = (2r + 1)/S - 1, mapped to the
sphere by the inverse Rosca-Lambert projection; the southern half (columns
S + c) holds the same (X, Y) with z -> -z. This is the convention the
bilinear lookup (grid_sample, align_corners=False) reads back. Intensity is
sum_b a_b exp(-asin(|n_b . x|)^2 / (2 sigma^2)) over band normals n_b, each
band family expanded over the proper rotations of the Laue group so the
pattern has the crystal's symmetry; then min-max normalised
(master_pattern.py:126-130).
"""

from __future__ import annotations

import math
from typing import Optional

import torch

from . import kernels
from .orientations import laue_elements, qu_apply, so3_uniform_quat


def inv_rosca_lambert(sq: torch.Tensor) -> torch.Tensor:
    """[-1, 1]^2 square -> northern unit hemisphere (Rosca, A&A 520, A63
    (2010)), float64."""
    L = math.sqrt(math.pi / 2)
    a = sq[..., 0] * L
    b = sq[..., 1] * L
    swap = b.abs() > a.abs()
    big = torch.where(swap, b, a)
    small = torch.where(swap, a, b)
    safe = torch.where(big == 0, torch.ones_like(big), big)
    r = (2 * big / math.pi) * torch.sqrt(torch.clamp(math.pi - big * big, min=0.0))
    ang = math.pi * small / (4 * safe)
    u = r * torch.cos(ang)
    v = r * torch.sin(ang)
    x = torch.where(swap, v, u)
    y = torch.where(swap, u, v)
    z = 1 - 2 * big * big / math.pi
    x = torch.where(big == 0, torch.zeros_like(x), x)
    y = torch.where(big == 0, torch.zeros_like(y), y)
    return torch.stack([x, y, z], -1)


def band_master_pattern(size: int, laue_id: int = 11, n_families: int = 40, sigma_deg: float = 2.0,
                        seed: int = 0, device: torch.device = torch.device("cpu")) -> torch.Tensor:
    """(size, 2*size) float32 Laue-symmetric band master pattern in [0, 1]."""
    g = torch.Generator().manual_seed(seed)
    normals = torch.randn(n_families, 3, generator=g, dtype=torch.float64)
    normals = normals / normals.norm(dim=-1, keepdim=True)
    amps = 0.5 + 0.5 * torch.rand(n_families, generator=g, dtype=torch.float64)
    ops = laue_elements(laue_id)
    # every family's normal under every symmetry rotation
    nrm = qu_apply(ops[:, None, :], normals[None, :, :]).reshape(-1, 3).to(device)
    amp = amps[None, :].expand(ops.shape[0], -1).reshape(-1).to(device)
    X = (2 * torch.arange(size, dtype=torch.float64, device=device) + 1) / size - 1
    yy, xx = torch.meshgrid(X, X, indexing="ij")
    nh = inv_rosca_lambert(torch.stack([xx, yy], -1)).reshape(-1, 3)
    sh = nh * torch.tensor([1.0, 1.0, -1.0], dtype=torch.float64, device=device)
    inv2s2 = 1.0 / (2 * math.radians(sigma_deg) ** 2)

    def intensity(p):
        out = torch.empty(p.shape[0], dtype=torch.float64, device=device)
        step = 1 << 16
        for s in range(0, p.shape[0], step):
            d = (p[s:s + step] @ nrm.T).abs().clamp(max=1.0)
            out[s:s + step] = (amp * torch.exp(-torch.asin(d) ** 2 * inv2s2)).sum(-1)
        return out

    mp = torch.cat([intensity(nh).view(size, size), intensity(sh).view(size, size)], dim=1)
    mp = (mp - mp.min()) / (1e-4 + mp.max() - mp.min())
    return mp.to(torch.float32)


def simulate_patterns(master_pattern: torch.Tensor, rays: torch.Tensor, orientations: torch.Tensor,
                      mean_counts=200.0, gaussian_sigma: float = 0.05, seed: int = 3,
                      chunk: int = 1 << 16) -> torch.Tensor:
    """Noisy patterns (N, D) float32 on the GPU: clean projection p, then
    counts ~ Poisson(mean_counts * p / mean(p)), out = counts / mean_counts +
    N(0, gaussian_sigma^2) (the counter-based noise kernel, keyed by seed and
    counted by chunk, row and pixel). mean_counts may be a per-pattern (N,)
    tensor."""
    dev = orientations.device
    n, d = orientations.shape[0], rays.shape[0]
    out = torch.empty((n, d), dtype=torch.float32, device=dev)
    mc = torch.as_tensor(mean_counts, dtype=torch.float32, device=dev)
    if mc.dim() == 0:
        mc = mc.expand(n)
    mc = mc.contiguous()
    quads = kernels.mp_quad_table(master_pattern)
    for c, s in enumerate(range(0, n, chunk)):
        q = orientations[s:s + chunk]
        clean = kernels.project(master_pattern, rays, q, zero_mean=False, quads=quads)
        kernels.poisson_gaussian_noise(clean.data, d, mc[s:s + chunk].contiguous(), gaussian_sigma, seed=seed,
                                       offset=c, out=out[s:s + chunk])
    return out


def random_orientations(n: int, seed: int = 2, device: torch.device = torch.device("cpu")) -> torch.Tensor:
    """Seeded uniform random orientations (w >= 0)."""
    g = torch.Generator().manual_seed(seed)
    q = so3_uniform_quat(n, generator=g).to(device)
    return torch.where(q[:, :1] < 0, -q, q).contiguous()
