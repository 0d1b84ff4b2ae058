"""Pattern-centre fitting by spatial cross-correlation (ebsd/fitting.py).

`spatial_fft_pc_fitter` mirrors fitting.py:359-535 (metric "dot"): the
detector is extended by a periphery that covers the searched pattern-centre
range (fitting.py:219-290), and for each pc_z of a sweep every orientation of
a coarse dictionary is projected onto that extended detector
(ebsd_project, no mean subtraction: the reference's raw
MasterPattern.interpolate) and cross-correlated with every zero-padded
experimental pattern through 2D FFTs (rocFFT through torch.fft); the best
dictionary score per shift is kept. The shift and pc_z of the largest
pattern-averaged score give the pattern centre.

Two properties of the reference are kept as they are: the orientation
reported for each pattern is quats_dict[i] with i the winner's index *within
its dictionary batch* (fitting.py:500-507, :528-529), which is the winning
orientation only when that falls in the first batch; and the centre is
derived from the best shift as 0.5 + (shift - padding - (n + 1) // 2) / n
(fitting.py:524-525), which for patterns simulated with the detector
geometry of geometry.py:31-92 at centre (x, y, z) gives (1 - x, 1 - y, z).
"""

from __future__ import annotations

from typing import Tuple

import torch

from . import kernels
from ._lib import DTYPE_F32
from .master_pattern import MasterPattern
from .orientations import sample_ori_fz_laue_angle


def bruker_geometries_to_SE3(pattern_centers: torch.Tensor, primary_tilt_deg: torch.Tensor,
                             secondary_tilt_deg: torch.Tensor, detector_shape: Tuple[int, int]):
    """fitting.py:53-114: rotation matrices (..., 3, 3) and translations
    (..., 3) of Bruker pattern centres, same expressions."""
    pcx, pcy, pcz = torch.unbind(pattern_centers, dim=-1)
    rows, cols = float(detector_shape[0]), float(detector_shape[1])
    a = torch.deg2rad(primary_tilt_deg)
    sx = torch.deg2rad(secondary_tilt_deg)
    rot = torch.stack([-torch.sin(a), -torch.sin(sx) * torch.cos(a), torch.cos(sx) * torch.cos(a),
                       torch.zeros_like(sx), torch.cos(sx), torch.sin(sx),
                       -torch.cos(a), torch.sin(sx) * torch.sin(a), -torch.sin(a) * torch.cos(sx)], dim=-1).view(-1, 3, 3)
    tx = (pcx * cols * torch.sin(sx) * torch.cos(a) + pcy * rows * torch.sin(a)
          + pcz * rows * torch.cos(sx) * torch.cos(a) - torch.sin(sx) * torch.cos(a) / 2 - torch.sin(a) / 2)
    ty = -cols * pcx * torch.cos(sx) + pcz * rows * torch.sin(sx) + torch.cos(sx) / 2
    tz = (-cols * pcx * torch.sin(sx) * torch.sin(a) + pcy * rows * torch.cos(a)
          - pcz * rows * torch.sin(a) * torch.cos(sx) + torch.sin(sx) * torch.sin(a) / 2 - torch.cos(a) / 2)
    return rot, torch.stack([tx, ty, tz], dim=-1)


def get_3D_pixel_coords_with_periphery(detector_shape: Tuple[int, int], device: torch.device, pc_x_min: float,
                                       pc_x_max: float, pc_y_min: float, pc_y_max: float, binning: int = 1,
                                       dtype: torch.dtype = torch.float32):
    """fitting.py:219-290: detector-frame pixel coordinates (N, 3) of the
    detector extended by the periphery, the padding (w_prior, w_after,
    h_prior, h_after) in binned pixels, and the extended binned shape."""
    b = float(binning)
    w0 = int((0.5 - pc_x_min) * detector_shape[1] / b)
    w1 = int((pc_x_max - 0.5) * detector_shape[1] / b)
    h0 = int((0.5 - pc_y_min) * detector_shape[0] / b)
    h1 = int((pc_y_max - 0.5) * detector_shape[0] / b)
    shape = (detector_shape[0] // binning + h0 + h1, detector_shape[1] // binning + w0 + w1)
    ii = torch.linspace(0.5 * (b - 1) - h0, detector_shape[0] + h1 - 0.5 * (b - 1), shape[0], dtype=dtype,
                        device=device)
    jj = torch.linspace(0.5 * (b - 1) - w0, detector_shape[1] + w1 - 0.5 * (b - 1), shape[1], dtype=dtype,
                        device=device)
    pix = torch.stack(torch.meshgrid(ii, jj, indexing="ij"), dim=-1).reshape(-1, 2)
    pix = torch.cat([pix, torch.zeros_like(pix[..., :1])], dim=-1)
    return pix, (w0, w1, h0, h1), shape


def periphery_rays(pixel_coords: torch.Tensor, rotation: torch.Tensor, translation: torch.Tensor) -> torch.Tensor:
    """Unit sample-frame rays of detector pixels (fitting.py:345-351)."""
    c = torch.matmul(rotation, pixel_coords[..., None]).squeeze(-1) + translation
    return c / torch.norm(c, dim=-1, keepdim=True)


def _levels(x: torch.Tensor, bins: int = 4) -> torch.Tensor:
    """(N, 1, H, W) -> (N, bins, H, W) float32 indicators of trunc(bins x) == k
    (utils/mutual_information.py:375-389, the uint8 cast of bins * x)."""
    lv = (bins * x).to(torch.uint8)
    return (lv == torch.arange(bins, dtype=torch.uint8, device=x.device)[None, :, None, None]).to(torch.float32)


def _mi_maps(sim: torch.Tensor, src_fft: torch.Tensor, pshape, chunk: int = 128) -> torch.Tensor:
    """cmif(images_dst=sim, images_src=patterns, src_ind=detector area, 4 + 4
    bins) for dictionary patterns sim (D, 1, Hp, Wp): (D, B, Hp, Wp). The
    25 cross-correlations of the indicator and level channels are FFT
    products; ebsd_cmif_maps turns them into entropies per shift. The
    inverse transform is taken at the padded shape (the reference's cmif
    leaves irfft2 its default size, one column short for odd widths, which
    its fitter cannot then combine; equal for even widths)."""
    d = sim.shape[0]
    out = []
    for s0 in range(0, d, chunk):
        x = sim[s0:s0 + chunk]
        dst = torch.cat((torch.ones_like(x), _levels(x)), 1)
        dst_fft = torch.fft.rfft2(dst)  # (d, 5, Hp, Wf)
        prod = dst_fft[:, None, :, None] * src_fft[None, :, None, :]  # (d, B, 5, 5, Hp, Wf)
        corr = torch.fft.irfft2(prod, s=tuple(pshape))
        del prod
        out.append(torch.fft.fftshift(kernels.cmif_maps(corr), dim=(-2, -1)))
        del corr
    return torch.cat(out, 0)


def spatial_fft_pc_fitter(mp: MasterPattern, patterns: torch.Tensor, pc_x_min: float = 0.0, pc_x_max: float = 1.0,
                          pc_y_min: float = 0.0, pc_y_max: float = 1.0, pc_z_min: float = 0.3,
                          pc_z_max: float = 0.7, pc_z_steps: int = 16, primary_tilt_deg: float = 70.0,
                          secondary_tilt_deg: float = 0.0, dictionary_resolution_deg: float = 1.0,
                          dictionary_indexing_batch_size: int = 1024, virtual_binning: int = 1,
                          metric: str = "dot", quats_dict: torch.Tensor = None):
    """fitting.py:359-535 on the GPU. patterns (B, 1, H, W) float32
    (modified in place, as the reference does). quats_dict: the dictionary
    (default: sample_ori_fz_laue_angle at dictionary_resolution_deg for the
    master pattern's Laue group). metric "dot": FFT cross-correlation of the
    zero-mean patterns; "mi": the mutual-information map of cmif with 4 + 4
    bins (fitting.py:486-496): level-indicator images of each dictionary
    pattern (all foreground) and each minmax-normalised experimental pattern
    (its detector area as foreground), their 25 channel-pair
    cross-correlations by FFT, and ebsd_cmif_maps for the entropies per
    shift. Returns (pattern centre (3,), orientations (B, 4))."""
    if metric not in ("dot", "mi"):
        raise ValueError("Invalid metric.")
    if patterns.dim() != 4:
        raise ValueError("Pattern must be 4D (B, C, H, W).")
    B, C, H, W = patterns.shape
    if C != 1:
        raise ValueError("Only single channel patterns are supported.")
    dev = patterns.device
    if dev.type != "cuda":
        raise ValueError("the fitter runs on the GPU: patterns must be on a CUDA device")
    hb, wb = H // virtual_binning, W // virtual_binning
    if hb * virtual_binning != H or wb * virtual_binning != W:
        raise ValueError("H and W must be divisible by the virtual binning.")
    if virtual_binning > 1:
        patterns = torch.nn.functional.avg_pool2d(patterns, kernel_size=virtual_binning, stride=virtual_binning)
    patterns -= patterns.view(B, -1).min(dim=-1, keepdim=True).values.view(B, 1, 1, 1)
    patterns /= patterns.view(B, -1).max(dim=-1, keepdim=True).values.view(B, 1, 1, 1)
    if metric == "dot":
        patterns -= patterns.view(B, -1).mean(dim=-1, keepdim=True).view(B, 1, 1, 1)

    pix, padding, pshape = get_3D_pixel_coords_with_periphery((H, W), dev, pc_x_min, pc_x_max, pc_y_min, pc_y_max,
                                                              binning=virtual_binning)
    pad = torch.nn.functional.pad(patterns, padding, mode="constant", value=0.0).view(1, B, pshape[0], pshape[1])
    if metric == "dot":
        pattern_fft = torch.fft.rfft2(pad).conj()
    else:
        ind = torch.nn.functional.pad(torch.ones((1, 1, hb, wb), device=dev), padding, mode="constant",
                                      value=0.0).expand(B, 1, pshape[0], pshape[1])
        src_fft = torch.fft.rfft2(torch.cat((ind, _levels(pad.view(B, 1, pshape[0], pshape[1]))), 1)).conj()
        # (B, 5, Hp, Wf): conj(FFT) of the source indicator and level images
    if quats_dict is None:
        quats_dict = sample_ori_fz_laue_angle(mp.laue_group, dictionary_resolution_deg, dev)
    quats_dict = quats_dict.to(dev, torch.float32).contiguous()
    mpb = mp.pattern_for_binning(virtual_binning).to(dev)
    quads = kernels.mp_quad_table(mpb, torch.float32)
    pc_z_vals = torch.linspace(pc_z_min, pc_z_max, pc_z_steps, device=dev)
    best_dots = torch.zeros((pc_z_steps, B, pshape[0], pshape[1]), device=dev)
    best_ori = torch.zeros((pc_z_steps, B, pshape[0], pshape[1]), dtype=torch.long, device=dev)
    prim = torch.tensor(primary_tilt_deg, device=dev)
    sec = torch.tensor(secondary_tilt_deg, device=dev)
    for z_idx in range(pc_z_steps):
        pc = torch.stack([torch.tensor(0.5, device=dev), torch.tensor(0.5, device=dev), pc_z_vals[z_idx]])[None]
        rot, trans = bruker_geometries_to_SE3(pc, -1.0 * prim, sec, (H, W))
        rays = periphery_rays(pix, rot, trans).contiguous()
        for j in range(0, quats_dict.shape[0], dictionary_indexing_batch_size):
            q = quats_dict[j:j + dictionary_indexing_batch_size]
            sim = kernels.project(mpb, rays, q, DTYPE_F32, zero_mean=False, quads=quads)
            sim = sim.data[:, :rays.shape[0]].reshape(q.shape[0], 1, pshape[0], pshape[1])
            if metric == "dot":
                dots = torch.fft.fftshift(torch.fft.irfft2(torch.fft.rfft2(sim, dim=(-2, -1)) * pattern_fft,
                                                           s=pshape), dim=(-2, -1))
            else:
                dots = _mi_maps(sim, src_fft, pshape)
            vals, idx = torch.max(dots, dim=0)
            best_dots[z_idx] = torch.maximum(best_dots[z_idx], vals)
            best_ori[z_idx] = torch.where(best_dots[z_idx] == vals, idx, best_ori[z_idx])
    avg = best_dots.mean(dim=1)
    bz, by, bx = torch.unravel_index(avg.argmax(), avg.shape)
    best_pc_z = pc_z_vals[bz]
    best_pc_y = 0.5 + float(by - padding[2] - (hb + 1) // 2) / hb
    best_pc_x = 0.5 + float(bx - padding[0] - (wb + 1) // 2) / wb
    best_quats = quats_dict[best_ori[bz, :, by, bx]]
    return torch.tensor([best_pc_x, best_pc_y, float(best_pc_z)], device=dev), best_quats
