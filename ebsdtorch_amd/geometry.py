"""Detector geometry: Bruker pattern centre + tilts -> sample-frame rays.

Mirrors ebsdtorch's EBSDGeometry (ebsd/geometry.py:139-503): same
constructor arguments, same pixel grid (linspace over [0, H] and [0, W]
with H/W points, 'ij' order, binning by fractional pixel centres) and the
same Bruker -> SE(3) conversion (geometry.py:31-92). This is once-per-call
setup of D rays; it runs as float32 torch arithmetic on the geometry's
device.
"""

from __future__ import annotations

import math
from typing import Optional, Tuple

import torch


def bruker_geometry_to_SE3(pattern_centers: torch.Tensor, primary_tilt_deg: torch.Tensor,
                           secondary_tilt_deg: torch.Tensor, detector_shape: Tuple[int, int]):
    """Rotation (..., 3, 3) and translation (..., 3) taking detector pixel
    coordinates (i, j, 0) to the sample frame, for pattern centres
    (pcx, pcy, pcz) in Bruker units, primary tilt = detector tilt - sample
    tilt and secondary tilt = sample x tilt (geometry.py:31-92)."""
    pcx, pcy, pcz = pattern_centers.unbind(-1)
    rows, cols = float(detector_shape[0]), float(detector_shape[1])
    a = torch.deg2rad(primary_tilt_deg)
    b = torch.deg2rad(secondary_tilt_deg)
    sa, ca, sb, cb = torch.sin(a), torch.cos(a), torch.sin(b), torch.cos(b)
    rot = torch.stack([-sa, -sb * ca, cb * ca,
                       torch.zeros_like(b), cb, sb,
                       -ca, sb * sa, -sa * cb], -1).view(-1, 3, 3)
    tx = pcx * cols * sb * ca + pcy * rows * sa + pcz * rows * cb * ca - sb * ca / 2 - sa / 2
    ty = -cols * pcx * cb + pcz * rows * sb + cb / 2
    tz = -cols * pcx * sb * sa + pcy * rows * ca - pcz * rows * sa * cb + sb * sa / 2 - ca / 2
    return rot, torch.stack([tx, ty, tz], -1)


def se3_exp(vec: torch.Tensor):
    """SE(3) exponential of (rx, ry, rz, tx, ty, tz): rotation by the
    Rodrigues formula and translation V t. Exact identity at zero."""
    w, t = vec[..., :3].double(), vec[..., 3:].double()
    th = w.norm(dim=-1, keepdim=True)
    eye = torch.eye(3, dtype=torch.float64, device=vec.device).expand(w.shape[:-1] + (3, 3))
    K = torch.zeros(w.shape[:-1] + (3, 3), dtype=torch.float64, device=vec.device)
    K[..., 0, 1], K[..., 0, 2], K[..., 1, 2] = -w[..., 2], w[..., 1], -w[..., 0]
    K[..., 1, 0], K[..., 2, 0], K[..., 2, 1] = w[..., 2], -w[..., 1], w[..., 0]
    small = th < 1e-8
    ths = torch.where(small, torch.ones_like(th), th)
    A = torch.where(small, 1 - th ** 2 / 6, torch.sin(ths) / ths)[..., None]
    B = torch.where(small, 0.5 - th ** 2 / 24, (1 - torch.cos(ths)) / ths ** 2)[..., None]
    C = torch.where(small, 1.0 / 6 - th ** 2 / 120, (ths - torch.sin(ths)) / ths ** 3)[..., None]
    KK = K @ K
    R = eye + A * K + B * KK
    V = eye + B * K + C * KK
    return R, (V @ t[..., None])[..., 0]


class EBSDGeometry(torch.nn.Module):
    """Detector/sample geometry (ebsd/geometry.py:139-503).

    detector_shape (H, W); tilts_degrees (sample x tilt, sample y tilt,
    detector tilt); proj_center (pcx, pcy, pcz) in the Bruker convention;
    se3_vector optional (1, 6) tune transform applied before the pattern
    centre transform; opt_rots / opt_shifts mask its components.
    """

    def __init__(self, detector_shape: Tuple[int, int], tilts_degrees=(0.0, 70.0, 0.0),
                 proj_center=(0.5, 0.5, 0.5), se3_vector: Optional[torch.Tensor] = None,
                 opt_rots=(False, False, False), opt_shifts=(False, False, False)):
        super().__init__()
        self.detector_shape = tuple(int(v) for v in detector_shape)
        tilts = [torch.tensor(float(t)) for t in tilts_degrees]
        self.register_buffer("sample_x_tilt_deg", tilts[0])
        self.register_buffer("sample_y_tilt_deg", tilts[1])
        self.register_buffer("detector_tilt_deg", tilts[2])
        rot, tr = bruker_geometry_to_SE3(torch.tensor(proj_center, dtype=torch.float32).view(1, 3),
                                         self.detector_tilt_deg - self.sample_y_tilt_deg,
                                         self.sample_x_tilt_deg, self.detector_shape)
        m = torch.zeros(1, 4, 4)
        m[..., :3, :3] = rot
        m[..., :3, 3] = tr
        m[..., 3, 3] = 1.0
        self.register_buffer("pc_SE3_matrix", m)
        self.se3_vector = torch.nn.Parameter(torch.zeros(1, 6) if se3_vector is None else se3_vector.view(1, 6))
        self.register_buffer("mask", torch.tensor(list(opt_rots) + list(opt_shifts), dtype=torch.bool))
        self.opt_se3 = bool(self.mask.any())

    def set_optimizable_se3_params(self, opt_rots=None, opt_shifts=None):
        if opt_rots is not None:
            self.mask[:3] = torch.tensor(opt_rots, dtype=torch.bool)
        if opt_shifts is not None:
            self.mask[3:] = torch.tensor(opt_shifts, dtype=torch.bool)
        self.opt_se3 = bool(self.mask.any())

    def get_detector2sample(self):
        """(3, 3) rotation and (3,) translation, detector -> sample frame."""
        vec = (self.se3_vector * self.mask).detach()
        if torch.count_nonzero(vec) == 0:
            m = self.pc_SE3_matrix[0]
        else:
            R, t = se3_exp(vec[0])
            tune = torch.eye(4, dtype=torch.float64, device=vec.device)
            tune[:3, :3], tune[:3, 3] = R, t
            m = (self.pc_SE3_matrix[0].double() @ tune).float()
        return m[:3, :3], m[:3, 3]

    def project2sample(self, pixel_coordinates: torch.Tensor, scan_coordinates=None) -> torch.Tensor:
        rot, tr = self.get_detector2sample()
        p = torch.cat([pixel_coordinates, torch.zeros_like(pixel_coordinates[..., :1])], -1)
        s = (rot @ p[..., None])[..., 0] + tr
        if scan_coordinates is not None:
            s = s - torch.cat([scan_coordinates, torch.zeros_like(scan_coordinates[..., :1])], -1)
        return s

    def backproject2detector(self, ray_directions: torch.Tensor, scan_coordinates=None) -> torch.Tensor:
        """Intersect sample-frame rays with the detector plane -> (i, j)."""
        rot, tr = self.get_detector2sample()
        origin = torch.zeros_like(tr) if scan_coordinates is None else torch.cat(
            [scan_coordinates, torch.zeros_like(scan_coordinates[..., :1])], -1)
        o = (origin - tr) @ rot
        tip = ray_directions if scan_coordinates is None else ray_directions + origin
        d = (tip - tr) @ rot - o
        t = -o[..., 2] / d[..., 2]
        return o[..., :2] + t[..., None] * d[..., :2]

    def get_coords_sample_frame(self, binning: Tuple[int, int], dtype=torch.float32) -> torch.Tensor:
        """(H/bh * W/bw, 3) un-normalised sample-frame pixel positions."""
        H, W = self.detector_shape
        if H % binning[0] != 0:
            raise ValueError(f"A height binning of {binning[0]} does not evenly divide the detector height of {H}.")
        if W % binning[1] != 0:
            raise ValueError(f"A width binning of {binning[1]} does not evenly divide the detector width of {W}.")
        b0, b1 = float(binning[0]), float(binning[1])
        dev = self.pc_SE3_matrix.device
        i = torch.linspace(0.5 * (b0 - 1), H - 0.5 * (b0 - 1), H // binning[0], dtype=dtype, device=dev)
        j = torch.linspace(0.5 * (b1 - 1), W - 0.5 * (b1 - 1), W // binning[1], dtype=dtype, device=dev)
        pix = torch.stack(torch.meshgrid(i, j, indexing="ij"), -1).view(-1, 2)
        return self.project2sample(pix)

    def pc_jacobian(self) -> torch.Tensor:
        """(3, 3) float32 J with: moving the pattern centre by dpc moves every
        un-normalised sample-frame pixel position by J @ dpc. The translation
        of bruker_geometry_to_SE3 is affine in the pattern centre and its
        rotation does not depend on it (nor does the se3 tune), so J is exact;
        it is read off as t(e_k) - t(0) in float64."""
        tilt = (self.detector_tilt_deg - self.sample_y_tilt_deg).double().cpu()
        sx = self.sample_x_tilt_deg.double().cpu()
        pcs = torch.cat([torch.zeros(1, 3, dtype=torch.float64), torch.eye(3, dtype=torch.float64)])
        _, t = bruker_geometry_to_SE3(pcs, tilt, sx, self.detector_shape)
        return (t[1:] - t[:1]).T.to(torch.float32).to(self.pc_SE3_matrix.device)

    def get_rays(self, binning: Tuple[int, int] = (1, 1)) -> torch.Tensor:
        """Unit sample-frame rays, (D, 3) float32 (indexing.py:96-106)."""
        c = self.get_coords_sample_frame(binning)
        return c / c.norm(dim=-1, keepdim=True)
