"""CPU oracle of ebsdtorch's spatial-FFT pattern-centre fitter
(ebsd/fitting.py:219-535, metric "dot").

TEST INFRASTRUCTURE ONLY. Only tests/ may import this module, and only as
the checker; the product package never imports it.

Restated with torch CPU float32 operations and the DI oracle's projection
(qu_apply + the side-by-side Lambert lookup): the periphery-extended pixel
grid, the per-pc_z detector rays, the FFT cross-correlation of every
dictionary pattern with every zero-padded pattern, the running best per
shift, and the final arg-max. Pinned in tests/test_oracle_fitting.py: the
periphery grid without periphery equals the detector grid (golden-pinned
geometry), the FFT correlation equals a scalar circular correlation, and a
synthetic scan's pattern centre is recovered.
"""

from __future__ import annotations

import math

import torch
import torch.nn.functional as F

from oracle import di_oracle as O


def periphery_grid(H: int, W: int, pc_x_min: float, pc_x_max: float, pc_y_min: float, pc_y_max: float,
                   binning: int = 1):
    """fitting.py:219-290."""
    b = float(binning)
    pad = (int((0.5 - pc_x_min) * W / b), int((pc_x_max - 0.5) * W / b),
           int((0.5 - pc_y_min) * H / b), int((pc_y_max - 0.5) * H / b))
    shape = (H // binning + pad[2] + pad[3], W // binning + pad[0] + pad[1])
    ii = torch.linspace(0.5 * (b - 1) - pad[2], H + pad[3] - 0.5 * (b - 1), shape[0])
    jj = torch.linspace(0.5 * (b - 1) - pad[0], W + pad[1] - 0.5 * (b - 1), shape[1])
    pix = torch.stack(torch.meshgrid(ii, jj, indexing="ij"), -1).reshape(-1, 2)
    return torch.cat((pix, torch.zeros(pix.shape[0], 1)), -1), pad, shape


def rays_at(pix: torch.Tensor, pc_z: float, H: int, W: int, primary_tilt_deg: float = 70.0,
            secondary_tilt_deg: float = 0.0) -> torch.Tensor:
    """fitting.py:456-461 + :345-351: rays of the (extended) detector at
    pattern centre (0.5, 0.5, pc_z), primary tilt negated as the fitter does."""
    R, t = O.bruker_geometry_to_SE3(torch.tensor([[0.5, 0.5, pc_z]]), torch.tensor(-primary_tilt_deg),
                                    torch.tensor(secondary_tilt_deg), (H, W))
    c = (R[0] @ pix[..., None]).squeeze(-1) + t[0]
    return c / c.norm(dim=-1, keepdim=True)


def correlate(sim: torch.Tensor, pats_fft_conj: torch.Tensor, shape) -> torch.Tensor:
    """fitting.py:490-492: fftshift(irfft2(rfft2(sim) * conj(rfft2(pat))))."""
    return torch.fft.fftshift(torch.fft.irfft2(torch.fft.rfft2(sim, dim=(-2, -1)) * pats_fft_conj, s=shape),
                              dim=(-2, -1))


def spatial_fft_pc_fitter(mp: torch.Tensor, quats_dict: torch.Tensor, patterns: torch.Tensor,
                          pc_x_min=0.0, pc_x_max=1.0, pc_y_min=0.0, pc_y_max=1.0, pc_z_min=0.3, pc_z_max=0.7,
                          pc_z_steps=16, primary_tilt_deg=70.0, secondary_tilt_deg=0.0, batch_size=1024):
    """fitting.py:359-535 (metric "dot", no binning) on the CPU. mp (S, 2S);
    patterns (B, 1, H, W). Returns (pattern centre (3,), orientations (B, 4))
    with the reference's batch-local orientation index."""
    B, _, H, W = patterns.shape
    p = patterns.clone()
    p -= p.view(B, -1).min(-1, keepdim=True).values.view(B, 1, 1, 1)
    p /= p.view(B, -1).max(-1, keepdim=True).values.view(B, 1, 1, 1)
    p -= p.view(B, -1).mean(-1, keepdim=True).view(B, 1, 1, 1)
    pix, pad, shape = periphery_grid(H, W, pc_x_min, pc_x_max, pc_y_min, pc_y_max)
    pf = torch.fft.rfft2(F.pad(p, pad).view(1, B, shape[0], shape[1])).conj()
    zs = torch.linspace(pc_z_min, pc_z_max, pc_z_steps)
    best = torch.zeros((pc_z_steps, B) + shape)
    best_i = torch.zeros((pc_z_steps, B) + shape, dtype=torch.long)
    for zi in range(pc_z_steps):
        rays = rays_at(pix, float(zs[zi]), H, W, primary_tilt_deg, secondary_tilt_deg)
        for j in range(0, quats_dict.shape[0], batch_size):
            q = quats_dict[j:j + batch_size]
            sim = O.interpolate(mp, O.qu_apply(q[:, None, :], rays[None]), normalize_coords=False)
            d = correlate(sim.view(q.shape[0], 1, shape[0], shape[1]), pf, shape)
            v, i = d.max(dim=0)
            best[zi] = torch.maximum(best[zi], v)
            best_i[zi] = torch.where(best[zi] == v, i, best_i[zi])
    avg = best.mean(1)
    bz, by, bx = torch.unravel_index(avg.argmax(), avg.shape)
    pc = torch.tensor([0.5 + float(bx - pad[0] - (W + 1) // 2) / W, 0.5 + float(by - pad[2] - (H + 1) // 2) / H,
                       float(zs[bz])])
    return pc, quats_dict[best_i[bz, :, by, bx]]
