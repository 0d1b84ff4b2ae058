"""CPU oracle of ebsdtorch's spatial-FFT pattern-centre fitter
(ebsd/fitting.py:219-535, metrics "dot" and "mi"), with the cross mutual
information map of utils/mutual_information.py:304-673 (cmif).

TEST INFRASTRUCTURE ONLY. Only tests/ may import this module, and only as
the checker; the product package never imports it.

Restated with torch CPU float32 operations and the DI oracle's projection
(qu_apply + the side-by-side Lambert lookup): the periphery-extended pixel
grid, the per-pc_z detector rays, the FFT cross-correlation of every
dictionary pattern with every zero-padded pattern, the running best per
shift, and the final arg-max. Pinned in tests/test_oracle_fitting.py: the
periphery grid without periphery equals the detector grid (golden-pinned
geometry), the FFT correlation equals a scalar circular correlation, a
synthetic scan's pattern centre is recovered, and cmif equals the
reference's own output on a fixture (tests/golden, make_golden.py) and a
brute-force histogram count per shift.
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


def cmif(images_dst: torch.Tensor, images_src: torch.Tensor, src_ind: torch.Tensor, bins_dst: int = 4,
         bins_src: int = 4) -> torch.Tensor:
    """mutual_information.py:304-673 with dst_ind=None, bins <= 8,
    constant_marginal=False, norm_mode "none": for destination images
    (Nd, ..., H, W) and source images (Ns, ..., H, W), the mutual information
    at every circular shift, (Nd, Ns, H, W)."""
    h, w = images_src.shape[-2:]
    d = images_dst.reshape(-1, 1, h, w)
    s = images_src.reshape(-1, 1, h, w)
    dl = (bins_dst * d).byte().repeat(1, bins_dst, 1, 1) == torch.arange(bins_dst, dtype=torch.uint8)[None, :, None, None]
    sl = (bins_src * s).byte().repeat(1, bins_src, 1, 1) == torch.arange(bins_src, dtype=torch.uint8)[None, :, None, None]
    dl = torch.cat((torch.ones_like(d), dl), 1)
    sl = torch.cat((src_ind.reshape(-1, 1, h, w), sl), 1)
    df = torch.fft.rfft2(dl.float())
    sf = torch.fft.rfft2(sl.float()).conj()
    cc = torch.fft.fftshift(torch.fft.irfft2(torch.einsum("bihw,cjhw->bcijhw", df, sf)), dim=(-2, -1)).real
    cc[cc < 1.0] = 0.0
    joint = cc[:, :, 1:, 1:]
    m_src = cc[:, :, 1:, 0]
    m_dst = cc[:, :, 0, 1:]
    joint = joint / joint.sum(dim=(2, 3), keepdim=True)
    m_src = m_src / m_src.sum(dim=2, keepdim=True)
    m_dst = m_dst / m_dst.sum(dim=2, keepdim=True)

    def plogp(x):
        return torch.where(x > 0.0, x * torch.log(x), torch.zeros_like(x))

    mi = -plogp(m_dst).sum(2) - plogp(m_src).sum(2) + plogp(joint).sum(dim=(2, 3))
    mi[torch.isnan(mi) | torch.isinf(mi) | (mi < 0.0)] = 0.0
    return mi


def cmif_brute(images_dst: torch.Tensor, images_src: torch.Tensor, src_ind: torch.Tensor, bins: int = 4):
    """cmif by direct counting: for each circular shift (fftshift order),
    the joint and marginal histograms of the overlapping pixel pairs counted
    one by one, in float64. Returns ((Nd, Ns, H, W) mutual information,
    (Nd, Ns, H, W) bool: shifts where some bin holds exactly one count;
    there the FFT route's round-off decides whether that count survives
    cmif's "below 1 -> 0" rule, so those shifts are not comparable)."""
    h, w = images_src.shape[-2:]
    d = images_dst.reshape(-1, h, w)
    s = images_src.reshape(-1, h, w)
    ind = src_ind.reshape(-1, h, w)
    dq = (bins * d).byte().long()
    sq = (bins * s).byte().long()
    out = torch.zeros((d.shape[0], s.shape[0], h, w), dtype=torch.float64)
    ones = torch.zeros((d.shape[0], s.shape[0], h, w), dtype=torch.bool)
    for a in range(d.shape[0]):
        for b in range(s.shape[0]):
            for y in range(h):
                for x in range(w):
                    # fftshift(irfft2(F(dst) conj(F(src))))[y, x] pairs dst[p + t] with src[p], t = (y, x) - centre
                    ty, tx = y - h // 2, x - w // 2
                    dd = torch.roll(dq[a], shifts=(-ty, -tx), dims=(0, 1))
                    joint = torch.zeros((bins, bins), dtype=torch.float64)
                    msrc = torch.zeros(bins, dtype=torch.float64)
                    mdst = torch.zeros(bins, dtype=torch.float64)
                    for i in range(bins):
                        di = (dd == i)
                        msrc[i] = (di.double() * ind[b].double()).sum()
                        mdst[i] = (sq[b] == i).double().sum()
                        for j in range(bins):
                            joint[i, j] = (di & (sq[b] == j)).double().sum()

                    def ent(hh):
                        hh = hh[hh >= 1.0]
                        if hh.numel() == 0:
                            return 0.0
                        p = hh / hh.sum()
                        return float(-(p * torch.log(p)).sum())

                    mi = ent(mdst) + ent(msrc) - ent(joint.reshape(-1))
                    out[a, b, y, x] = max(mi, 0.0)
                    ones[a, b, y, x] = bool((joint == 1).any() or (msrc == 1).any() or (mdst == 1).any())
    return out, ones


def spatial_fft_pc_fitter(mp: torch.Tensor, quats_dict: torch.Tensor, patterns: torch.Tensor,
                          pc_x_min=0.0, pc_x_max=1.0, pc_y_min=0.0, pc_y_max=1.0, pc_z_min=0.3, pc_z_max=0.7,
                          pc_z_steps=16, primary_tilt_deg=70.0, secondary_tilt_deg=0.0, batch_size=1024,
                          metric: str = "dot"):
    """fitting.py:359-535 (no binning) on the CPU. mp (S, 2S); patterns
    (B, 1, H, W). Returns (pattern centre (3,), orientations (B, 4)) with the
    reference's batch-local orientation index."""
    B, _, H, W = patterns.shape
    p = patterns.clone()
    p -= p.view(B, -1).min(-1, keepdim=True).values.view(B, 1, 1, 1)
    p /= p.view(B, -1).max(-1, keepdim=True).values.view(B, 1, 1, 1)
    if metric == "dot":
        p -= p.view(B, -1).mean(-1, keepdim=True).view(B, 1, 1, 1)
    pix, pad, shape = periphery_grid(H, W, pc_x_min, pc_x_max, pc_y_min, pc_y_max)
    ppad = F.pad(p, pad).view(1, B, shape[0], shape[1])
    pf = torch.fft.rfft2(ppad).conj()
    ind = F.pad(torch.ones((1, 1, H, W)), pad).repeat(B, 1, 1, 1)
    zs = torch.linspace(pc_z_min, pc_z_max, pc_z_steps)
    best = torch.zeros((pc_z_steps, B) + shape)
    best_i = torch.zeros((pc_z_steps, B) + shape, dtype=torch.long)
    for zi in range(pc_z_steps):
        rays = rays_at(pix, float(zs[zi]), H, W, primary_tilt_deg, secondary_tilt_deg)
        for j in range(0, quats_dict.shape[0], batch_size):
            q = quats_dict[j:j + batch_size]
            sim = O.interpolate(mp, O.qu_apply(q[:, None, :], rays[None]), normalize_coords=False)
            sim = sim.view(q.shape[0], 1, shape[0], shape[1])
            if metric == "dot":
                d = correlate(sim, pf, shape)
            else:
                d = cmif(sim, ppad, ind)
            v, i = d.max(dim=0)
            best[zi] = torch.maximum(best[zi], v)
            best_i[zi] = torch.where(best[zi] == v, i, best_i[zi])
    avg = best.mean(1)
    bz, by, bx = torch.unravel_index(avg.argmax(), avg.shape)
    pc = torch.tensor([0.5 + float(bx - pad[0] - (W + 1) // 2) / W, 0.5 + float(by - pad[2] - (H + 1) // 2) / H,
                       float(zs[bz])])
    return pc, quats_dict[best_i[bz, :, by, bx]]
