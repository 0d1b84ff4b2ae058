"""CPU oracle of ebsdtorch's dictionary-indexing (DI) path.

TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker (or, in the
bench, as the timed CPU baseline); the product package never imports it.

Every function restates one reference function with the same torch CPU
operations in the same order, so on the CPU it reproduces the reference's
float32 results. Citations are file:line in ZacharyVarley/ebsdtorch v0.0.64.

How each part is pinned (tests/test_oracle_*.py):
  * sampler, FZ tests, Laue tables ..... reference golden vectors
    (tests/golden/ref_vectors.npz, written by tests/golden/make_golden.py
    from the reference modules in the build container) + invariants
    (unit norm, w >= 0, brute-force FZ, group closure, counts)
  * detector geometry ................. golden vectors + back-projection
    round trip + closed form at the pattern centre
  * qu_apply ........................... golden vectors + rotation-matrix
    identity in float64
  * side-by-side Lambert ............... golden vectors + special points +
    inverse round trip + equal-area histogram
  * grid_sample gather ................. golden vectors + scalar brute force
  * ChunkedKNN streaming top-k ......... golden vectors + float64 brute force
    + invariance to chunk sizes
  * OnlineCovMatrix .................... golden vectors + numpy.cov (float64)
  * end-to-end DI (small) .............. golden vectors of the reference's
    own dictionary_index_orientations (indexing.py:14-211)
"""

from __future__ import annotations

import math
from typing import Optional, Tuple

import torch
import torch.nn.functional as F

# --------------------------------------------------------------------------
# Laue groups (s2_and_so3/laue_generators.py:14-249). The operator tables are
# data of the reference (Larsen & Schmidt 2017); their order matters only in
# that the identity comes first (ori_in_fz_laue_brute relies on it).
# --------------------------------------------------------------------------

_R2 = 1.0 / (2.0 ** 0.5)
_R3 = (3.0 ** 0.5) / 2.0

_LAUE_TABLES = {
    1: [[1, 0, 0, 0]],
    2: [[1, 0, 0, 0], [0, 0, 0, 1]],
    3: [[1, 0, 0, 0], [0, 0, 0, 1], [0, 1, 0, 0], [0, 0, 1, 0]],
    4: [[1, 0, 0, 0], [0, 0, 0, 1], [_R2, 0, 0, _R2], [_R2, 0, 0, -_R2]],
    5: [[1, 0, 0, 0], [0, 0, 0, 1], [0, 1, 0, 0], [0, 0, 1, 0],
        [_R2, 0, 0, _R2], [_R2, 0, 0, -_R2], [0, _R2, _R2, 0], [0, -_R2, _R2, 0]],
    6: [[1, 0, 0, 0], [0.5, 0, 0, _R3], [0.5, 0, 0, -_R3]],
    7: [[1, 0, 0, 0], [0.5, 0, 0, _R3], [0.5, 0, 0, -_R3],
        [0, 1, 0, 0], [0, -0.5, _R3, 0], [0, 0.5, _R3, 0]],
    8: [[1, 0, 0, 0], [0.5, 0, 0, _R3], [0.5, 0, 0, -_R3],
        [0, 0, 0, 1], [_R3, 0, 0, 0.5], [_R3, 0, 0, -0.5]],
    9: [[1, 0, 0, 0], [0.5, 0, 0, _R3], [0.5, 0, 0, -_R3], [0, 0, 0, 1],
        [_R3, 0, 0, 0.5], [_R3, 0, 0, -0.5], [0, 1, 0, 0], [0, -0.5, _R3, 0],
        [0, 0.5, _R3, 0], [0, _R3, 0.5, 0], [0, -_R3, 0.5, 0], [0, 0, 1, 0]],
    10: [[1, 0, 0, 0], [0, 0, 0, 1], [0, 1, 0, 0], [0, 0, 1, 0],
         [0.5, 0.5, -0.5, 0.5], [0.5, 0.5, 0.5, -0.5], [0.5, 0.5, -0.5, -0.5],
         [0.5, -0.5, -0.5, -0.5], [0.5, -0.5, 0.5, 0.5], [0.5, -0.5, 0.5, -0.5],
         [0.5, -0.5, -0.5, 0.5], [0.5, 0.5, 0.5, 0.5]],
    11: [[1, 0, 0, 0], [0, 0, 0, 1], [0, 1, 0, 0], [0, 0, 1, 0],
         [_R2, 0, 0, _R2], [_R2, 0, 0, -_R2], [0, _R2, _R2, 0], [0, -_R2, _R2, 0],
         [0.5, 0.5, -0.5, 0.5], [0.5, 0.5, 0.5, -0.5], [0.5, 0.5, -0.5, -0.5],
         [0.5, -0.5, -0.5, -0.5], [0.5, -0.5, 0.5, 0.5], [0.5, -0.5, 0.5, -0.5],
         [0.5, -0.5, -0.5, 0.5], [0.5, 0.5, 0.5, 0.5], [_R2, _R2, 0, 0],
         [_R2, -_R2, 0, 0], [_R2, 0, _R2, 0], [_R2, 0, -_R2, 0], [0, _R2, 0, _R2],
         [0, -_R2, 0, _R2], [0, 0, _R2, _R2], [0, 0, -_R2, _R2]],
}

_LAUE_MULTS = [2, 4, 8, 8, 16, 6, 12, 12, 24, 24, 48]


def get_laue_mult(laue_id: int) -> int:
    """laue_generators.py:14-51."""
    return _LAUE_MULTS[laue_id - 1]


def laue_elements(laue_id: int) -> torch.Tensor:
    """laue_generators.py:54-249 (identity first). The reference builds its
    float64 tables inside TorchScript, which rounds the literals through
    float32 first; the same rounding is applied here (pinned by the golden
    vectors)."""
    return torch.tensor(_LAUE_TABLES[laue_id], dtype=torch.float32).to(torch.float64)


# --------------------------------------------------------------------------
# Quaternions (s2_and_so3/quaternions.py)
# --------------------------------------------------------------------------


def qu_std(qu: torch.Tensor) -> torch.Tensor:
    """quaternions.py:27-38."""
    return torch.where(qu[..., 0:1] >= 0, qu, -qu)


def qu_prod(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Hamilton product standardised to w >= 0 (quaternions.py:55-92:
    qu_prod_raw + qu_std, same expression order)."""
    aw, ax, ay, az = a[..., 0], a[..., 1], a[..., 2], a[..., 3]
    bw, bx, by, bz = b[..., 0], b[..., 1], b[..., 2], b[..., 3]
    ow = aw * bw - ax * bx - ay * by - az * bz
    ox = aw * bx + ax * bw + ay * bz - az * by
    oy = aw * by - ax * bz + ay * bw + az * bx
    oz = aw * bz + ax * by - ay * bx + az * bw
    return qu_std(torch.stack((ow, ox, oy, oz), -1))


def qu_conj(qu: torch.Tensor) -> torch.Tensor:
    """quaternions.py:171-183."""
    return qu * torch.tensor([1, -1, -1, -1], dtype=qu.dtype)


def qu_apply(qu: torch.Tensor, point: torch.Tensor) -> torch.Tensor:
    """quaternions.py:186-213, same expression order (q (1, p) q*)."""
    aw, ax, ay, az = qu[..., 0], qu[..., 1], qu[..., 2], qu[..., 3]
    bx, by, bz = point[..., 0], point[..., 1], point[..., 2]
    iw = aw - ax * bx - ay * by - az * bz
    ix = aw * bx + ax + ay * bz - az * by
    iy = aw * by - ax * bz + ay + az * bx
    iz = aw * bz + ax * by - ay * bx + az
    ox = -iw * ax + ix * aw - iy * az + iz * ay
    oy = -iw * ay + ix * az + iy * aw - iz * ax
    oz = -iw * az - ix * ay + iy * ax + iz * aw
    return torch.stack((ox, oy, oz), -1)


def qu2om(qu: torch.Tensor) -> torch.Tensor:
    """Rotation matrix of a unit quaternion (orientations.py:911-937
    semantics: active rotation, v' = R v)."""
    w, x, y, z = torch.unbind(qu, -1)
    r = torch.stack(
        [
            1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
            2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
            2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y),
        ],
        -1,
    )
    return r.reshape(qu.shape[:-1] + (3, 3))


# --------------------------------------------------------------------------
# Cubochoric sampler (orientations.py:293-349, :419-516, :730-746,
# :1136-1147; sampling.py:179-205; laue_fz_ori.py:144-239, :472-557)
# --------------------------------------------------------------------------


def cu2ho(cu: torch.Tensor) -> torch.Tensor:
    """orientations.py:293-349."""
    indices = torch.argsort(torch.abs(cu), dim=-1, descending=False)
    srt = torch.gather(cu, -1, indices)
    s, m, b = srt.unbind(dim=-1)
    trig_arg_xy = s * torch.pi / (12.0 * m)
    trig_arg_xy[torch.isnan(trig_arg_xy)] = 0
    factor_xy = (
        2 ** (1 / 12) * 3 ** (1 / 3) * m
        * torch.sqrt(
            (4 * b ** 2 * (torch.cos(trig_arg_xy) - (2 ** 0.5))
             + (2 ** 0.5) * m ** 2 * (-2 * (2 ** 0.5) * torch.cos(trig_arg_xy) + 3))
            / (torch.cos(trig_arg_xy) - (2 ** 0.5))
        )
        / (torch.pi ** (1 / 3) * b * torch.sqrt(-torch.cos(trig_arg_xy) + (2 ** 0.5)))
    )
    x_s3 = factor_xy * torch.sin(trig_arg_xy)
    y_s3 = factor_xy * 2 ** -0.5 * ((2 ** 0.5) * torch.cos(trig_arg_xy) - 1)
    z_s3 = (
        2 * 6 ** (1 / 3) * b ** 2 * (torch.cos(trig_arg_xy) - (2 ** 0.5))
        + 2 ** (5 / 6) * 3 ** (1 / 3) * m ** 2 * (-2 * (2 ** 0.5) * torch.cos(trig_arg_xy) + 3)
    ) / (2 * torch.pi ** (1 / 3) * b * (torch.cos(trig_arg_xy) - (2 ** 0.5)))
    ho = torch.stack((x_s3, y_s3, z_s3), dim=-1)
    ho = torch.scatter(ho, -1, indices, ho)
    ho[torch.isnan(ho)] = 0
    ho.copysign_(cu)
    return ho


_HO2AX_FIT_F32 = [
    1.0000000000000009e00, -4.9999943403867775e-01, -2.5015165060149020e-02,
    -3.8120131548551729e-03, -1.2106188330642162e-03, 4.9329295993155416e-04,
    -7.0089385526450620e-04, 3.0979774923589078e-04, -7.3023474963298843e-05,
]
_HO2AX_FIT_F64 = [
    1.0000000000000000e00, -4.9999997124013285e-01, -2.5001181866044025e-02,
    -3.9144209820521038e-03, -8.9320268104539483e-04, 3.1181024286083695e-05,
    -4.3961032788396477e-04, 3.9657471727506439e-04, -2.6379945050586932e-04,
    9.1185355979587159e-05, -1.4875867805692529e-05,
]


def ho2ax(ho: torch.Tensor) -> torch.Tensor:
    """orientations.py:419-516 (fast=True branch). The fit coefficients are
    the reference's published Chebyshev fit (data)."""
    coeffs = _HO2AX_FIT_F32 if ho.dtype in (torch.float32, torch.float16) else _HO2AX_FIT_F64
    fit = torch.tensor(coeffs, dtype=ho.dtype)
    ho_norm_sq = torch.sum(ho ** 2, dim=-1, keepdim=False)
    s = torch.zeros_like(ho_norm_sq)
    for i in range(len(fit)):
        s += fit[i] * ho_norm_sq ** i
    w = 2.0 * torch.arccos(torch.clamp(s, -1.0, 1.0))
    ax = torch.concat([ho * torch.rsqrt(ho_norm_sq).unsqueeze(-1), w.unsqueeze(-1)], dim=-1)
    rot_is_identity = torch.abs(ho_norm_sq) < 1e-6
    ax[rot_is_identity] = 0
    ax[rot_is_identity, ..., 2] = 1.0
    return ax


def ax2qu(ax: torch.Tensor) -> torch.Tensor:
    """orientations.py:730-746."""
    qu = torch.empty_like(ax)
    cos_half_ang = torch.cos(ax[..., 3] / 2.0)
    sin_half_ang = torch.sin(ax[..., 3:4] / 2.0)
    qu[..., 0] = cos_half_ang
    qu[..., 1:] = ax[..., :3] * sin_half_ang
    return qu


def cu2qu(cu: torch.Tensor) -> torch.Tensor:
    """orientations.py:1136-1147."""
    return ax2qu(ho2ax(cu2ho(cu)))


def so3_cubochoric_grid(edge_length: int) -> torch.Tensor:
    """sampling.py:179-205."""
    cu = torch.linspace(-0.5 * torch.pi ** (2 / 3), 0.5 * torch.pi ** (2 / 3), edge_length + 1)
    cu = cu[:-1]
    cu = torch.stack(torch.meshgrid(cu, cu, cu, indexing="ij"), dim=-1).reshape(-1, 3)
    return qu_std(cu2qu(cu))


def ori_in_fz_laue(quats: torch.Tensor, laue_id: int) -> torch.Tensor:
    """laue_fz_ori.py:144-239 (inclusive bounds)."""
    w = quats[..., 0]
    if laue_id == 11:
        xyz_abs = torch.abs(quats[..., 1:])
        return (torch.max(xyz_abs, dim=-1).values <= (w * (2 ** 0.5 - 1))) & (
            torch.sum(xyz_abs, dim=-1) <= w)
    if laue_id == 10:
        return torch.sum(torch.abs(quats[..., 1:]), dim=-1) <= w
    if laue_id == 9:
        x_abs, y_abs = torch.abs(quats[..., 1]), torch.abs(quats[..., 2])
        cond = x_abs > y_abs
        m = torch.where(cond, x_abs, y_abs)
        n = torch.where(cond, y_abs, x_abs)
        rot = torch.where(m > (2 + 3 ** 0.5) * n, m, (3 ** 0.5 / 2) * m + 0.5 * n)
        return (torch.abs(quats[..., 3]) <= (2 - 3 ** 0.5) * w) & (rot <= w)
    if laue_id == 8:
        return torch.abs(quats[..., 3]) <= ((2 - 3 ** 0.5) * w)
    if laue_id == 7:
        rot = torch.where(
            torch.abs(quats[..., 1]) >= torch.abs(quats[..., 2]) * (3 ** 0.5),
            torch.abs(quats[..., 1]),
            (3 ** 0.5 / 2) * torch.abs(quats[..., 2]) + 0.5 * torch.abs(quats[..., 1]),
        )
        return (torch.abs(quats[..., 3]) <= ((1.0 / 3 ** 0.5) * w)) & (rot <= w)
    if laue_id == 6:
        return torch.abs(quats[..., 3]) <= (1.0 / 3 ** 0.5) * w
    if laue_id == 5:
        x_abs, y_abs = torch.abs(quats[..., 1]), torch.abs(quats[..., 2])
        cond = x_abs > y_abs
        m = torch.where(cond, x_abs, y_abs)
        n = torch.where(cond, y_abs, x_abs)
        rot = torch.where(m > (2 ** 0.5 + 1) * n, m, (1 / 2 ** 0.5) * m + (1 / 2 ** 0.5) * n)
        return (torch.abs(quats[..., 3]) <= ((2 ** 0.5 - 1) * w)) & (rot <= w)
    if laue_id == 4:
        return torch.abs(quats[..., 3]) <= (2 ** 0.5 - 1) * w
    if laue_id == 3:
        return torch.max(torch.abs(quats[..., 1:]), dim=-1).values <= w
    if laue_id == 2:
        return torch.abs(quats[..., 3]) <= w
    if laue_id == 1:
        return torch.full(quats.shape[:-1], True, dtype=torch.bool)
    raise ValueError(f"Laue group {laue_id} is not supported")


def ori_in_fz_laue_brute(quats: torch.Tensor, laue_id: int) -> torch.Tensor:
    """laue_fz_ori.py:108-141: in the FZ iff the identity gives the largest
    |real part| among all symmetry-equivalent quaternions."""
    ops = laue_elements(laue_id).to(quats.dtype)
    eq = qu_prod(quats.reshape(-1, 1, 4), ops.reshape(1, -1, 4))
    return (torch.argmax(eq[..., 0].abs(), dim=-1) == 0).reshape(quats.shape[:-1])


def sample_ori_fz_laue(laue_id: int, target_n_samples: int, permute_seed: Optional[int] = None) -> torch.Tensor:
    """laue_fz_ori.py:472-513. The reference permutes with an unseeded
    randperm; here the permutation is optional and seeded."""
    edge_length = int((target_n_samples * 0.5 * get_laue_mult(laue_id)) ** (1.0 / 3.0))
    s = so3_cubochoric_grid(edge_length)
    s = s[ori_in_fz_laue(s, laue_id)]
    if permute_seed is not None:
        g = torch.Generator().manual_seed(permute_seed)
        s = s[torch.randperm(s.shape[0], generator=g)]
    return s


def sample_ori_fz_laue_angle(laue_id: int, angular_resolution_deg: float,
                             permute_seed: Optional[int] = None) -> torch.Tensor:
    """laue_fz_ori.py:516-557."""
    n = (2 * 131.97049 / (angular_resolution_deg - 0.03732) + 1) ** 3
    edge_length = int((n / (0.5 * get_laue_mult(laue_id))) ** (1.0 / 3.0) + 1.0)
    s = so3_cubochoric_grid(edge_length)
    s = s[ori_in_fz_laue(s, laue_id)]
    if permute_seed is not None:
        g = torch.Generator().manual_seed(permute_seed)
        s = s[torch.randperm(s.shape[0], generator=g)]
    return s


def disori_angle_laue(q1: torch.Tensor, q2: torch.Tensor, laue_id: int) -> torch.Tensor:
    """laue_fz_ori.py:414-469 (same Laue group on both sides): radians."""
    mis = qu_prod(q1, qu_conj(q2))
    ops = laue_elements(laue_id).to(q1.dtype)
    # |real part of g2 * mis * g1| over all operator pairs
    left = qu_prod(ops.reshape(1, -1, 1, 4), mis.reshape(-1, 1, 1, 4))
    eq = qu_prod(left, ops.reshape(1, 1, -1, 4))
    c = eq[..., 0].abs().reshape(mis.shape[0], -1).max(dim=-1).values
    return 2.0 * torch.acos(torch.clamp(c, -1.0, 1.0))


# --------------------------------------------------------------------------
# Detector geometry (ebsd/geometry.py:31-92, :186-231, :257-347, :422-503)
# --------------------------------------------------------------------------


def bruker_geometry_to_SE3(pattern_centers: torch.Tensor, primary_tilt_deg: torch.Tensor,
                           secondary_tilt_deg: torch.Tensor, detector_shape: Tuple[int, int]):
    """geometry.py:31-92."""
    pcx, pcy, pcz = torch.unbind(pattern_centers, dim=-1)
    rows, cols = float(detector_shape[0]), float(detector_shape[1])
    dt_m_sy = torch.deg2rad(primary_tilt_deg)
    sx = torch.deg2rad(secondary_tilt_deg)
    rot = torch.stack(
        [
            -torch.sin(dt_m_sy), -torch.sin(sx) * torch.cos(dt_m_sy), torch.cos(sx) * torch.cos(dt_m_sy),
            torch.zeros_like(sx), torch.cos(sx), torch.sin(sx),
            -torch.cos(dt_m_sy), torch.sin(sx) * torch.sin(dt_m_sy), -torch.sin(dt_m_sy) * torch.cos(sx),
        ],
        dim=-1,
    ).view(-1, 3, 3)
    tx = (pcx * cols * torch.sin(sx) * torch.cos(dt_m_sy) + pcy * rows * torch.sin(dt_m_sy)
          + pcz * rows * torch.cos(sx) * torch.cos(dt_m_sy)
          - torch.sin(sx) * torch.cos(dt_m_sy) / 2 - torch.sin(dt_m_sy) / 2)
    ty = -cols * pcx * torch.cos(sx) + pcz * rows * torch.sin(sx) + torch.cos(sx) / 2
    tz = (-cols * pcx * torch.sin(sx) * torch.sin(dt_m_sy) + pcy * rows * torch.cos(dt_m_sy)
          - pcz * rows * torch.sin(dt_m_sy) * torch.cos(sx)
          + torch.sin(sx) * torch.sin(dt_m_sy) / 2 - torch.cos(dt_m_sy) / 2)
    return rot, torch.stack([tx, ty, tz], dim=-1)


def detector_to_sample_transform(detector_shape, tilts_degrees=(0.0, 70.0, 0.0), proj_center=(0.5, 0.5, 0.5)):
    """EBSDGeometry.__init__ + get_detector2sample with a zero SE(3) tune
    vector (geometry.py:186-231, :257-280). se3_exp_map_om(0) is the exact
    identity (lie_algebra/se3.py:187-205), so the transform is the pattern
    centre SE(3) matrix itself."""
    t = tuple(torch.tensor(v) for v in tilts_degrees)
    rot, tr = bruker_geometry_to_SE3(torch.tensor(proj_center).view(1, 3), t[2] - t[1], t[0], detector_shape)
    m = torch.zeros(1, 4, 4)
    m[..., :3, :3] = rot
    m[..., :3, 3] = tr
    m[..., 3, 3] = 1.0
    m = m @ torch.eye(4)[None]
    return m[..., :3, :3], m[..., :3, 3]


def detector_coords(detector_shape, tilts_degrees=(0.0, 70.0, 0.0), proj_center=(0.5, 0.5, 0.5),
                    binning=(1, 1)) -> torch.Tensor:
    """EBSDGeometry.get_coords_sample_frame (geometry.py:422-503) ->
    project2sample (:282-347). Un-normalised sample-frame coordinates (D, 3)."""
    H, W = detector_shape
    if H % binning[0] != 0 or W % binning[1] != 0:
        raise ValueError("binning must divide the detector shape")
    bh, bw = H // binning[0], W // binning[1]
    b0, b1 = float(binning[0]), float(binning[1])
    pix = torch.stack(
        torch.meshgrid(
            torch.linspace(0.5 * (b0 - 1), H - (0.5 * (b0 - 1)), bh, dtype=torch.float32),
            torch.linspace(0.5 * (b1 - 1), W - (0.5 * (b1 - 1)), bw, dtype=torch.float32),
            indexing="ij",
        ),
        dim=-1,
    ).view(-1, 2)
    rot, tr = detector_to_sample_transform(detector_shape, tilts_degrees, proj_center)
    pix = torch.cat([pix, torch.zeros_like(pix[..., 0:1])], dim=-1)
    return ((rot @ pix[..., None]) + tr[..., None]).squeeze(-1)


def backproject_to_detector(rays: torch.Tensor, detector_shape, tilts_degrees=(0.0, 70.0, 0.0),
                            proj_center=(0.5, 0.5, 0.5)) -> torch.Tensor:
    """EBSDGeometry.backproject2detector without scan offsets
    (geometry.py:349-420): sample-frame ray -> detector pixel (i, j)."""
    rot, tr = detector_to_sample_transform(detector_shape, tilts_degrees, proj_center)
    origin = -tr @ rot
    tips = (rays - tr) @ rot
    d = tips - origin
    t = -origin[..., 2] / d[..., 2]
    return origin[..., :2] + t[..., None] * d[..., :2]


def normalized_rays(detector_shape, proj_center, tilts_degrees=(0.0, 70.0, 0.0), binning=(1, 1)):
    """indexing.py:96-106 (average pattern centre)."""
    c = detector_coords(detector_shape, tilts_degrees, proj_center, binning)
    return c / c.norm(dim=-1, keepdim=True)


# --------------------------------------------------------------------------
# Lambert projection and master-pattern sampling
# --------------------------------------------------------------------------


def rosca_lambert_side_by_side(pts: torch.Tensor) -> torch.Tensor:
    """sphere.py:145-185."""
    x, y, z = pts[..., 0], pts[..., 1], pts[..., 2]
    factor = torch.sqrt(torch.clamp(2.0 * (1.0 - torch.abs(z)), min=0.0))
    cond = torch.abs(y) <= torch.abs(x)
    big = torch.where(cond, x, y)
    sml = torch.where(cond, y, x)
    simpler_term = torch.where(big < 0, -1, 1) * factor * (2.0 / (8.0 ** 0.5))
    arctan_term = (
        torch.where(big < 0, -1, 1) * factor
        * torch.atan2(sml * torch.where(big < 0, -1, 1), torch.abs(big))
        * (2.0 * (2.0 ** 0.5) / torch.pi)
    )
    out = torch.stack((simpler_term, arctan_term), dim=-1)
    out = torch.where(cond[..., None], out, out.flip(-1))
    out[..., 0] = (out[..., 0] / 2.0) - 0.5 + torch.where(z < 0, 1.0, 0)
    return out


def inv_rosca_lambert(pts: torch.Tensor) -> torch.Tensor:
    """sphere.py:95-142: square [-1, 1]^2 -> northern hemisphere."""
    pi = torch.pi
    x_abs = torch.abs(pts[..., 0]) * (pi / 2) ** 0.5
    y_abs = torch.abs(pts[..., 1]) * (pi / 2) ** 0.5
    cond = x_abs >= y_abs
    x_new = torch.where(cond, x_abs, y_abs)
    y_new = torch.where(cond, y_abs, x_abs)
    x_hs = (2 * x_new / pi) * torch.sqrt(pi - x_new ** 2) * torch.cos(pi * y_new / (4 * x_new))
    y_hs = (2 * x_new / pi) * torch.sqrt(pi - x_new ** 2) * torch.sin(pi * y_new / (4 * x_new))
    z_out = 1 - (2 * x_new ** 2 / pi)
    x_out = torch.where(cond, x_hs, y_hs)
    y_out = torch.where(cond, y_hs, x_hs)
    x_out = x_out.copysign_(pts[..., 0])
    y_out = y_out.copysign_(pts[..., 1])
    return torch.stack((x_out, y_out, z_out), dim=-1)


def interpolate(mp: torch.Tensor, coords_3d: torch.Tensor, normalize_coords: bool = True,
                mode: str = "bilinear", padding_mode: str = "border", align_corners: bool = False) -> torch.Tensor:
    """MasterPattern.interpolate (master_pattern.py:141-209): side-by-side
    Lambert coordinates sampled by grid_sample with the given mode, padding
    and corner alignment (defaults: the dictionary-indexing lookup). The
    near-square reshape of :186-195 only changes the grid's shape, not the
    values, so a (1, 1, n, 2) grid is used."""
    if normalize_coords:
        coords_3d = coords_3d / torch.norm(coords_3d, dim=-1, keepdim=True)
    g = rosca_lambert_side_by_side(coords_3d)
    n = g.numel() // 2
    out = F.grid_sample(mp[None, None], g.reshape(1, 1, n, 2), mode=mode,
                        padding_mode=padding_mode, align_corners=align_corners)
    return out.view(coords_3d.shape[:-1])


def project_dictionary(mp: torch.Tensor, rays: torch.Tensor, quats: torch.Tensor,
                       chunk: int = 256, normalize_coords: bool = False,
                       signal_mask: Optional[torch.Tensor] = None) -> torch.Tensor:
    """indexing.py:113-143: rotate, interpolate, mask, zero-mean (N, D)."""
    outs = []
    for qb in torch.split(quats, chunk):
        rot = qu_apply(qb[:, None, :], rays[None, ...])
        sim = interpolate(mp, rot, normalize_coords=normalize_coords).view(len(qb), -1)
        if signal_mask is not None:
            sim = sim[:, signal_mask.flatten()]
        outs.append(sim - torch.mean(sim, dim=-1, keepdim=True))
    return torch.cat(outs, 0)


# --------------------------------------------------------------------------
# Streaming top-k (utils/knn.py:80-287), angular metric, float32, no quant
# --------------------------------------------------------------------------


class _NoBiasLinear(torch.nn.Module):
    """The bias-free Linear layer holding a dictionary chunk as its weight
    (knn.py:36-49)."""

    def __init__(self, weight: torch.Tensor):
        super().__init__()
        self.fc = torch.nn.Linear(weight.shape[1], weight.shape[0], bias=False)
        self.fc.weight.data = weight

    def forward(self, x):
        return self.fc(x)


def quant_model_data(data: torch.Tensor) -> torch.nn.Module:
    """knn.py:52-58: the chunk as the weight of a Linear, dynamically
    quantised to int8 (per-tensor weight scale; activations quantised per call)
    by torch's quantize_dynamic (the library call the reference makes)."""
    from torch.ao.quantization import quantize_dynamic

    layer = _NoBiasLinear(data)
    layer.eval()
    return quantize_dynamic(model=layer, qconfig_spec={torch.nn.Linear}, dtype=torch.qint8, inplace=True)


def topk_quantized_data(query_batch: torch.Tensor, quantized_data: torch.nn.Module, k: int):
    """knn.py:70-72."""
    res = torch.topk(quantized_data(query_batch), k, dim=1, largest=True, sorted=False)
    return res.values, res.indices


class ChunkedKNN:
    """utils/knn.py:106-287, angular metric. match_dtype float32; quantized
    selects the reference's CPU default (quantized_via_ao=True on a CPU
    device: int8 dynamic quantisation of each data chunk, knn.py:135-146,
    :177-180, :196-199, :236-239)."""

    def __init__(self, data_size: int, query_size: int, topk: int, match_dtype=torch.float32,
                 quantized: bool = False):
        self.data_size, self.query_size, self.topk = data_size, query_size, topk
        self.quantized = bool(quantized)
        if self.quantized:
            match_dtype = torch.float32  # knn.py:135-143
        self.match_dtype = match_dtype
        ind_dtype = torch.int64 if data_size < 2 ** 31 else torch.int32
        self.knn_indices = torch.empty((query_size, topk), dtype=ind_dtype)
        self.knn_distances = torch.full((query_size, topk), -torch.inf, dtype=match_dtype)
        self.prepared_data = None
        self.curr_end = 0

    def set_data_chunk(self, data_chunk: torch.Tensor):
        """knn.py:163-180."""
        data_chunk = data_chunk.to(self.match_dtype)
        self.curr_end += data_chunk.shape[0]
        self.curr_start = self.curr_end - data_chunk.shape[0]
        self.prepared_data = quant_model_data(data_chunk) if self.quantized else data_chunk

    def _search(self, query):
        if self.quantized:
            return topk_quantized_data(query, self.prepared_data, self.topk)
        return self._knn_batch(self.prepared_data, query, self.topk)

    @staticmethod
    def _knn_batch(data, query, topk):
        """knn.py:80-103, angular."""
        dist = query @ data.t()
        return torch.topk(dist, topk, dim=1, largest=True, sorted=False)

    def query_all(self, query: torch.Tensor):
        """knn.py:182-217."""
        if self.prepared_data is None:
            raise ValueError("Data chunk is not set.")
        d, i = self._search(query.to(self.match_dtype))
        i = i + self.curr_start
        md = torch.cat((self.knn_distances, d), dim=1)
        mi = torch.cat((self.knn_indices, i), dim=1)
        top = torch.topk(md, self.topk, dim=1, largest=True, sorted=False)[1]
        self.knn_indices = torch.gather(mi, 1, top)
        self.knn_distances = torch.gather(md, 1, top)

    def query_chunk(self, query_chunk: torch.Tensor, query_start: int):
        """knn.py:219-266."""
        if self.prepared_data is None:
            raise ValueError("Data chunk is not set.")
        query_end = query_start + query_chunk.shape[0]
        d, i = self._search(query_chunk.to(self.match_dtype))
        i = i + self.curr_start
        md = torch.cat((self.knn_distances[query_start:query_end], d), dim=1)
        mi = torch.cat((self.knn_indices[query_start:query_end], i), dim=1)
        top = torch.topk(md, self.topk, dim=1, largest=True, sorted=False)[1]
        self.knn_indices[query_start:query_end] = torch.gather(mi, 1, top)
        self.knn_distances[query_start:query_end] = torch.gather(md, 1, top)

    def retrieve_topk(self):
        """knn.py:268-287."""
        top = torch.topk(self.knn_distances, self.topk, dim=1, largest=True, sorted=True)[1]
        return torch.gather(self.knn_indices, 1, top), torch.gather(self.knn_distances, 1, top)


def dictionary_index(mp: torch.Tensor, rays: torch.Tensor, exp_pats: torch.Tensor, dictionary: torch.Tensor,
                     top_k: int = 1, dictionary_chunk_size: int = 256, experiment_chunk_size: int = 256,
                     signal_mask: Optional[torch.Tensor] = None, match_dtype=torch.float32,
                     quantized: bool = False):
    """dictionary_index_orientations (indexing.py:14-211) for one phase, with
    the dictionary passed in (the reference draws it with an unseeded
    randperm); plain FP32 matching, or with quantized=True the reference's
    CPU default (quantized_if_x86, int8 quantize_dynamic of each dictionary
    chunk). exp_pats: (N, H*W) float32.
    Returns (indices (N, k) int64, scores (N, k))."""
    n_exp = exp_pats.shape[0]
    knn = ChunkedKNN(len(dictionary), n_exp, top_k, match_dtype, quantized=quantized)
    for ori_batch in torch.split(dictionary, dictionary_chunk_size):
        rot = qu_apply(ori_batch[:, None, :], rays[None, ...])
        sim = interpolate(mp, rot, normalize_coords=False).view(len(ori_batch), -1)
        if signal_mask is not None:
            sim = sim[:, signal_mask.flatten()]
        sim = sim - torch.mean(sim, dim=-1, keepdim=True)
        knn.set_data_chunk(sim)
        if experiment_chunk_size > n_exp:
            q = exp_pats
            if signal_mask is not None:
                q = q[:, signal_mask.flatten()]
            q = q - torch.mean(q, dim=-1, keepdim=True)
            knn.query_all(q)
        else:
            for idx in torch.split(torch.arange(n_exp), experiment_chunk_size):
                q = exp_pats[idx].view(len(idx), -1)
                if signal_mask is not None:
                    q = q[:, signal_mask.flatten()]
                q = q - torch.mean(q, dim=-1, keepdim=True)
                knn.query_chunk(q, query_start=int(idx[0]))
    return knn.retrieve_topk()


# --------------------------------------------------------------------------
# Streaming covariance and PCA DI (utils/pca_covmat.py:53-158,
# indexing.py:214-476)
# --------------------------------------------------------------------------


class OnlineCovMatrix:
    """pca_covmat.py:94-158 (Welford/Chan batched update)."""

    def __init__(self, n_features: int, covmat_dtype=torch.float32, delta_dtype=torch.float32):
        self.covmat_dtype, self.delta_dtype = covmat_dtype, delta_dtype
        self.mean = torch.zeros(1, n_features, dtype=covmat_dtype)
        self.covmat_aggregate = torch.zeros((n_features, n_features), dtype=covmat_dtype)
        self.obs = torch.tensor([0], dtype=torch.int64)

    def update(self, data_new: torch.Tensor):
        """update_covmat, pca_covmat.py:53-91."""
        n = data_new.shape[0]
        batch_mean = torch.mean(data_new, dim=0, keepdim=True)
        new_mean = (self.mean * self.obs + batch_mean * n) / (self.obs + n)
        delta = data_new.to(self.delta_dtype) - self.mean.to(self.delta_dtype)
        delta_prime = data_new.to(self.delta_dtype) - new_mean.to(self.delta_dtype)
        self.covmat_aggregate += torch.einsum("ij,ik->jk", delta, delta_prime).to(self.covmat_dtype)
        self.obs += n
        self.mean.copy_(new_mean)

    def get_covmat(self):
        """pca_covmat.py:133-147 (correlation=False)."""
        return self.covmat_aggregate / (self.obs - 1).to(self.covmat_dtype)

    def get_eigenvectors(self):
        """pca_covmat.py:149-158: ascending eigenvalues."""
        return torch.linalg.eigh(self.get_covmat())[1]


def pca_components(eigenvectors: torch.Tensor, n_components: int, reference_slice: bool = False):
    """indexing.py:355. With reference_slice the reference's literal slice
    [:, -n : -n + 300] is applied; otherwise the top-n components (the
    documented intent) are returned, largest eigenvalue last."""
    if reference_slice:
        return eigenvectors[:, -n_components:(-n_components + 300)]
    return eigenvectors[:, -n_components:]


def pca_dictionary_index(mp, coords, exp_pats, learn_dictionary, index_dictionary, n_components, top_k=1,
                         dictionary_chunk_size=256, experiment_chunk_size=256, covmat_dtype=torch.float32,
                         reference_slice=False):
    """pca_dictionary_index_orientations (indexing.py:214-476), one phase,
    explicit dictionaries, FP32 matching. `coords` are the un-normalised
    detector coordinates (indexing.py:312-317); interpolate normalises them
    after rotation (normalize_coords=True, :329-335)."""
    n_exp = exp_pats.shape[0]
    cov = OnlineCovMatrix(coords.shape[0], covmat_dtype=covmat_dtype, delta_dtype=covmat_dtype)
    for ori_batch in torch.split(learn_dictionary, dictionary_chunk_size):
        rot = qu_apply(ori_batch[:, None, :], coords[None, ...])
        sim = interpolate(mp, rot, normalize_coords=True).view(len(ori_batch), -1)
        sim = sim - torch.mean(sim, dim=-1, keepdim=True)
        cov.update(sim.to(covmat_dtype))
    evecs = torch.linalg.eigh(cov.get_covmat())[1]
    pca = pca_components(evecs, n_components, reference_slice).to(torch.float32)
    knn = ChunkedKNN(len(index_dictionary), n_exp, top_k)
    for ori_batch in torch.split(index_dictionary, dictionary_chunk_size):
        rot = qu_apply(ori_batch[:, None, :], coords[None, ...])
        sim = interpolate(mp, rot, normalize_coords=True).view(len(ori_batch), -1)
        sim = sim - torch.mean(sim, dim=-1, keepdim=True)
        knn.set_data_chunk(torch.matmul(sim, pca))
        for idx in torch.split(torch.arange(n_exp), experiment_chunk_size):
            q = exp_pats[idx].view(len(idx), -1)
            q = q - torch.mean(q, dim=-1, keepdim=True)
            knn.query_chunk(torch.matmul(q, pca), query_start=int(idx[0]))
    return knn.retrieve_topk()


# --------------------------------------------------------------------------
# Orientation grid refinement (ebsd/refining.py:13-244), one phase, average
# pattern centre
# --------------------------------------------------------------------------


def get_local_orientation_grid(semi_edge_in_degrees: float, kernel_radius_in_steps: int,
                               axial_grid_dimension: int = 3) -> torch.Tensor:
    """refining.py:13-68 (cubochoric coordinates, float32)."""
    semi = semi_edge_in_degrees * torch.pi ** (2 / 3) / 360.0
    cu = torch.linspace(-semi, semi, 2 * kernel_radius_in_steps + 1)
    cu = torch.stack(torch.meshgrid(cu, cu, cu, indexing="ij"), dim=-1).reshape(-1, 3)
    if axial_grid_dimension == 2:
        cu = cu[torch.any(cu == 0, dim=-1)]
    elif axial_grid_dimension == 1:
        cu = cu[torch.sum(cu == 0, dim=-1) >= 2]
    elif axial_grid_dimension != 3:
        raise ValueError("axial_grid_dimension must be 1, 2, or 3.")
    return cu


def ori_to_fz_laue(quats: torch.Tensor, laue_id: int) -> torch.Tensor:
    """laue_fz_ori.py:35-81: the right product with the symmetry operator
    that maximises |w| (qu_prod_pos_real, argmax: first maximum)."""
    shape = quats.shape
    ops = laue_elements(laue_id).to(quats.dtype)
    q = quats.reshape(-1, 1, 4)
    w = (q[..., 0] * ops[:, 0] - q[..., 1] * ops[:, 1] - q[..., 2] * ops[:, 2] - q[..., 3] * ops[:, 3]).abs()
    best = torch.argmax(w, dim=-1)
    return qu_prod(quats.reshape(-1, 4), ops[best]).reshape(shape)


def interpolate_aligned(mp: torch.Tensor, coords_3d: torch.Tensor) -> torch.Tensor:
    """MasterPattern.interpolate with mode bilinear, padding border,
    align_corners=True, normalize_coords=False (master_pattern.py:141-209, as
    refining.py:198-204 calls it)."""
    g = rosca_lambert_side_by_side(coords_3d)
    n = g.numel() // 2
    out = F.grid_sample(mp[None, None], g.reshape(1, 1, n, 2), mode="bilinear",
                        padding_mode="border", align_corners=True)
    return out.view(coords_3d.shape[:-1])


def orientation_grid_refinement(mp: torch.Tensor, laue_id: int, rays: torch.Tensor, exp_pats: torch.Tensor,
                                orientations: torch.Tensor, grid_semi_edge_in_degrees: float, batch_size: int,
                                virtual_binning: int = 1, n_iter: int = 3, axial_grid_dimension: int = 3,
                                kernel_radius_in_steps: int = 1, shrink_factor: float = 0.5,
                                match_dtype=torch.float32):
    """orientation_grid_refinement (refining.py:71-244) for one phase with
    an average pattern centre. mp (S, 2S) unbinned master pattern; rays
    (D_b, 3) unit sample-frame directions of the binned detector
    (get_coords_sample_frame(binning=(b, b)) normalised, refining.py:112-129);
    exp_pats (N, H, W) unbinned patterns; orientations (N, 4).
    Returns (refined orientations (N, 4), dots (N, n_iter))."""
    mp_b = mp
    if virtual_binning > 1:
        # master_pattern.py:94-113 (MasterPattern.bin)
        mp_b = F.avg_pool2d(mp[None, None], virtual_binning, stride=virtual_binning)[0, 0]
    n = exp_pats.shape[0]
    cu_grid = get_local_orientation_grid(grid_semi_edge_in_degrees, kernel_radius_in_steps,
                                         axial_grid_dimension).view(1, -1, 3)
    dots = torch.zeros((n, n_iter))
    out = orientations.clone()
    for idx in torch.split(torch.arange(n), batch_size):
        cu_cur = cu_grid.clone().repeat(len(idx), 1, 1)
        pats = exp_pats[idx]
        if virtual_binning > 1:
            # experiment_pats.py:342
            pats = F.avg_pool2d(pats[:, None], virtual_binning)[:, 0]
        pats = pats.reshape(len(idx), -1)
        pats = pats - torch.mean(pats, dim=-1, keepdim=True)
        q_cur = out[idx]
        for j in range(n_iter):
            q_grid = cu2qu(cu_cur)
            q_aug = qu_prod(q_grid, q_cur[:, None, :])
            q_aug = ori_to_fz_laue(q_aug, laue_id)
            rot = qu_apply(q_aug[:, :, None, :], rays[None, None, :, :])
            sim = interpolate_aligned(mp_b, rot).view(len(idx), q_aug.shape[1], -1)
            sim = sim - torch.mean(sim, dim=-1, keepdim=True)
            d = torch.einsum("ij,ikj->ik", pats.to(match_dtype), sim.to(match_dtype))
            best_d, best_i = torch.max(d, dim=-1)
            dots[idx, j] = best_d.float()
            q_cur = q_aug[torch.arange(len(idx)), best_i]
            cu_cur[best_i == ((cu_cur.shape[1] - 1) / 2)] *= shrink_factor
        out[idx] = q_cur
    return out, dots


def orientation_grad_refinement(mp: torch.Tensor, laue_id: int, rays: torch.Tensor, exp_pats: torch.Tensor,
                                orientations: torch.Tensor, batch_size: int, virtual_binning: int = 1,
                                n_iter: int = 50, learning_rate: float = 0.001,
                                match_dtype=torch.float16) -> torch.Tensor:
    """orientation_grad_refinement (refining.py:247-388) for one phase with an
    average pattern centre: Adam on the quaternions, the loss minus the mean
    cosine between each zero-mean unit pattern and its bicubic
    (align_corners=True) simulation, back-propagated by autograd through
    grid_sample; after n_iter steps the quaternions go to the positive
    hemisphere and the fundamental zone, and are kept only where the final
    dot is below minus the first one (refining.py:381-388). mp (S, 2S)
    unbinned; rays (D_b, 3) unit sample-frame directions of the binned
    detector; exp_pats (N, H, W); orientations (N, 4). Returns (N, 4)."""
    mp_b = mp
    if virtual_binning > 1:
        mp_b = F.avg_pool2d(mp[None, None], virtual_binning, stride=virtual_binning)[0, 0]
    n = exp_pats.shape[0]
    out = orientations.clone()
    for idx in torch.split(torch.arange(n), batch_size):
        pats = exp_pats[idx]
        if virtual_binning > 1:
            pats = F.avg_pool2d(pats[:, None], virtual_binning)[:, 0]
        pats = pats.reshape(len(idx), -1)
        pats = pats - torch.mean(pats, dim=-1, keepdim=True)
        pats = pats / pats.norm(dim=-1, keepdim=True)
        q = out[idx].clone().detach().requires_grad_(True)
        opt = torch.optim.Adam([q], lr=learning_rate)
        initial = None
        for step in range(n_iter):
            rot = qu_apply(q[:, None, :], rays[None, :, :])
            sim = interpolate(mp_b, rot, normalize_coords=False, mode="bicubic", align_corners=True).squeeze()
            sim = sim - torch.mean(sim, dim=-1, keepdim=True)
            sim = sim / sim.norm(dim=-1, keepdim=True)
            d = torch.einsum("ij,ij->i", pats.to(match_dtype), sim.to(match_dtype))
            if step == 0:
                initial = -d
            loss = -torch.mean(d)
            opt.zero_grad()
            loss.backward()
            opt.step()
            with torch.no_grad():
                q /= q.norm(dim=-1, keepdim=True)
        q = q.detach()
        q *= (q[..., :1] >= 0).float() * 2 - 1
        q = ori_to_fz_laue(q, laue_id)
        upd = d < initial
        out[idx[upd]] = q[upd]
    return out
