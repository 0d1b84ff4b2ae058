"""Orientation dictionary sampler and the SO(3) helpers the DI path needs.

Public names follow ebsdtorch (s2_and_so3/laue_fz_ori.py,
s2_and_so3/sampling.py, s2_and_so3/laue_generators.py,
s2_and_so3/quaternions.py). Quaternions are (w, x, y, z) with w >= 0.

The cubochoric grid is mapped to quaternions in float64 on the target device
(cube -> ball by the Rosca-Morawiec-De Graef equal-volume map, ball ->
rotation angle by Newton iteration on (3/4)(w - sin w) = |h|^3) and rounded
once to float32, so every returned quaternion is within float32 rounding of
the exact grid point.
"""

from __future__ import annotations

import itertools
import math
from functools import lru_cache
from typing import Optional

import torch

LAUE_NAMES = {
    1: "C1", 2: "C2", 3: "D2", 4: "C4", 5: "D4", 6: "C3", 7: "D3", 8: "C6", 9: "D6", 10: "T", 11: "O",
}
# order of the proper rotation group times 2 (inversion), laue_generators.py:14-51
_LAUE_MULT = {1: 2, 2: 4, 3: 8, 4: 8, 5: 16, 6: 6, 7: 12, 8: 12, 9: 24, 10: 24, 11: 48}


def get_laue_mult(laue_id: int) -> int:
    """Multiplicity of Laue group `laue_id` (1..11), inversion included."""
    if laue_id not in _LAUE_MULT:
        raise ValueError(f"Laue group {laue_id} is not supported")
    return _LAUE_MULT[laue_id]


def _axis_angle_quat(axis, angle):
    a = torch.tensor(axis, dtype=torch.float64)
    a = a / a.norm()
    return torch.cat([torch.tensor([math.cos(angle / 2)], dtype=torch.float64), a * math.sin(angle / 2)])


def _generators(laue_id: int):
    """Rotation generators of each Laue group's proper subgroup, in the
    crystal frame used by ebsdtorch (c along z; hexagonal a1 along x)."""
    z2, z3, z4, z6 = ([0, 0, 1], math.pi), ([0, 0, 1], 2 * math.pi / 3), ([0, 0, 1], math.pi / 2), (
        [0, 0, 1], math.pi / 3)
    x2 = ([1, 0, 0], math.pi)
    d3 = ([1, 1, 1], 2 * math.pi / 3)
    return {
        1: [],
        2: [z2],
        3: [z2, x2],
        4: [z4],
        5: [z4, x2],
        6: [z3],
        7: [z3, x2],
        8: [z6],
        9: [z6, x2],
        10: [z2, x2, d3],
        11: [z4, x2, d3],
    }[laue_id]


def _qmul(a, b):
    aw, ax, ay, az = a
    bw, bx, by, bz = b
    return torch.stack([
        aw * bw - ax * bx - ay * by - az * bz,
        aw * bx + ax * bw + ay * bz - az * by,
        aw * by - ax * bz + ay * bw + az * bx,
        aw * bz + ax * by - ay * bx + az * bw,
    ])


def _canonical(q: torch.Tensor) -> torch.Tensor:
    """q and -q are one rotation: keep the sign that makes the first
    non-zero component positive."""
    for v in q:
        if abs(float(v)) > 1e-12:
            return q if v > 0 else -q
    return q


@lru_cache(maxsize=None)
def _laue_elements_cpu(laue_id: int) -> torch.Tensor:
    gens = [_axis_angle_quat(a, t) for a, t in _generators(laue_id)]
    elems = [torch.tensor([1.0, 0.0, 0.0, 0.0], dtype=torch.float64)]
    frontier = list(elems)
    while frontier:
        new = []
        for e, g in itertools.product(frontier, gens):
            p = _canonical(_qmul(g, e))
            if all((p - q).abs().max() > 1e-9 for q in elems):
                elems.append(p)
                new.append(p)
        frontier = new
    ident, rest = elems[0], elems[1:]
    rest.sort(key=lambda q: (-round(q[0].item(), 9), *[-round(v.item(), 9) for v in q[1:]]))
    out = torch.stack([ident, *rest])
    if out.shape[0] != get_laue_mult(laue_id) // 2:
        raise RuntimeError("Laue group closure failed")
    return out


def laue_elements(laue_id: int) -> torch.Tensor:
    """(|G|, 4) float64 unit quaternions of the proper rotations of Laue
    group `laue_id`; the identity comes first."""
    get_laue_mult(laue_id)
    return _laue_elements_cpu(laue_id).clone()


def qu_std(q: torch.Tensor) -> torch.Tensor:
    return torch.where(q[..., :1] >= 0, q, -q)


def qu_prod(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    aw, ax, ay, az = a.unbind(-1)
    bw, bx, by, bz = b.unbind(-1)
    return torch.stack([
        aw * bw - ax * bx - ay * by - az * bz,
        aw * bx + ax * bw + ay * bz - az * by,
        aw * by - ax * bz + ay * bw + az * bx,
        aw * bz + ax * by - ay * bx + az * bw,
    ], -1)


def qu_conj(q: torch.Tensor) -> torch.Tensor:
    return torch.cat([q[..., :1], -q[..., 1:]], -1)


def qu_apply(q: torch.Tensor, v: torch.Tensor) -> torch.Tensor:
    """Rotate points v (..., 3) by unit quaternions q (..., 4) (broadcast):
    the vector part of q (1, v) q*, expanded term by term in the rounding
    order of quaternions.py:186-213 (the projection kernel uses the same
    order), so float32 results equal the reference's bit for bit. The 1 in
    the scalar slot only adds |q|^2 to the scalar part, which is dropped."""
    aw, ax, ay, az = q.unbind(-1)
    bx, by, bz = v.unbind(-1)
    pw = aw - ax * bx - ay * by - az * bz
    px = aw * bx + ax + ay * bz - az * by
    py = aw * by - ax * bz + ay + az * bx
    pz = aw * bz + ax * by - ay * bx + az
    return torch.stack((-pw * ax + px * aw - py * az + pz * ay,
                        -pw * ay + px * az + py * aw - pz * ax,
                        -pw * az - px * ay + py * ax + pz * aw), -1)


# --------------------------------------------------------------------------
# cubochoric -> quaternion
# --------------------------------------------------------------------------

_A = math.pi ** (2.0 / 3.0)  # cube edge


def _cu2ho(cu: torch.Tensor) -> torch.Tensor:
    """Equal-volume cube -> ball map (Rosca, Morawiec & De Graef, MSMSE 22
    (2014) 075013), float64. Points are handled in the pyramid of their
    largest |coordinate|."""
    mag = cu.abs()
    order = torch.argsort(mag, dim=-1)
    s, m, b = torch.gather(cu, -1, order).abs().unbind(-1)
    with torch.no_grad():
        t = torch.where(m > 0, s * math.pi / (12.0 * m), torch.zeros_like(m))
        c = torch.cos(t)
        r2 = 2.0 ** 0.5
        num = 4 * b * b * (c - r2) + r2 * m * m * (3 - 2 * r2 * c)
        f = (2 ** (1 / 12) * 3 ** (1 / 3) * m * torch.sqrt(num / (c - r2))
             / (math.pi ** (1 / 3) * b * torch.sqrt(r2 - c)))
        x = f * torch.sin(t)
        y = f * (r2 * c - 1) / r2
        z = (2 * 6 ** (1 / 3) * b * b * (c - r2) + 2 ** (5 / 6) * 3 ** (1 / 3) * m * m * (3 - 2 * r2 * c)) / (
            2 * math.pi ** (1 / 3) * b * (c - r2))
        ho_sorted = torch.stack([x, y, z], -1)
        ho_sorted = torch.nan_to_num(ho_sorted, nan=0.0)
        ho = torch.empty_like(ho_sorted).scatter_(-1, order, ho_sorted)
        return torch.copysign(ho.abs(), cu)


def _ho2qu(ho: torch.Tensor) -> torch.Tensor:
    """Ball -> quaternion: |h|^3 = (3/4)(w - sin w), solved for w by Newton
    iteration in float64 (converges to machine precision on [0, pi])."""
    h2 = (ho * ho).sum(-1)
    h = torch.sqrt(h2)
    target = h2 * h  # |h|^3
    # initial guess from the small-angle series w^3/8 = |h|^3 -> w = 2|h|,
    # clipped into the domain
    w = torch.clamp(2.0 * h, max=math.pi)
    for _ in range(40):
        f = 0.75 * (w - torch.sin(w)) - target
        fp = 0.75 * (1 - torch.cos(w))
        step = torch.where(fp > 1e-300, f / fp, torch.zeros_like(f))
        w = torch.clamp(w - step, 0.0, math.pi)
    axis = torch.where(h[..., None] > 0, ho / h.clamp_min(1e-300)[..., None], torch.zeros_like(ho))
    return torch.cat([torch.cos(w / 2)[..., None], axis * torch.sin(w / 2)[..., None]], -1)


def cu2qu(cu: torch.Tensor) -> torch.Tensor:
    """Cubochoric (..., 3) -> unit quaternions (..., 4), w >= 0, float32."""
    q = _ho2qu(_cu2ho(cu.to(torch.float64)))
    return qu_std(q).to(torch.float32)


def so3_cubochoric_grid(edge_length: int, device: torch.device = torch.device("cpu")) -> torch.Tensor:
    """edge_length^3 cubochoric grid points (half-open cube) as quaternions
    (sampling.py:179-205 grid layout, 'ij' order)."""
    lin = torch.linspace(-0.5 * _A, 0.5 * _A, edge_length + 1, dtype=torch.float64, device=device)[:-1]
    cu = torch.stack(torch.meshgrid(lin, lin, lin, indexing="ij"), -1).reshape(-1, 3)
    return cu2qu(cu)


def ori_in_fz_laue(q: torch.Tensor, laue_id: int) -> torch.Tensor:
    """Membership in the orientation fundamental zone (inclusive bounds),
    the closed forms of laue_fz_ori.py:144-239."""
    w, x, y, z = q.unbind(-1)
    ax, ay, az = x.abs(), y.abs(), z.abs()
    r2m1, r3 = 2 ** 0.5 - 1, 3 ** 0.5
    if laue_id == 11:
        return (torch.maximum(torch.maximum(ax, ay), az) <= w * r2m1) & (ax + ay + az <= w)
    if laue_id == 10:
        return ax + ay + az <= w
    if laue_id in (9, 5):
        big = torch.maximum(ax, ay)
        small = torch.minimum(ax, ay)
        if laue_id == 9:
            rot = torch.where(big > (2 + r3) * small, big, (r3 / 2) * big + 0.5 * small)
            return (az <= (2 - r3) * w) & (rot <= w)
        rot = torch.where(big > (2 ** 0.5 + 1) * small, big, (big + small) / 2 ** 0.5)
        return (az <= r2m1 * w) & (rot <= w)
    if laue_id == 8:
        return az <= (2 - r3) * w
    if laue_id == 7:
        rot = torch.where(ax >= ay * r3, ax, (r3 / 2) * ay + 0.5 * ax)
        return (az <= w / r3) & (rot <= w)
    if laue_id == 6:
        return az <= w / r3
    if laue_id == 4:
        return az <= r2m1 * w
    if laue_id == 3:
        return torch.maximum(torch.maximum(ax, ay), az) <= w
    if laue_id == 2:
        return az <= w
    if laue_id == 1:
        return torch.ones_like(w, dtype=torch.bool)
    raise ValueError(f"Laue group {laue_id} is not supported")


def _fz_grid(laue_id: int, edge: int, device, permute: bool, generator: Optional[torch.Generator]):
    # build the grid in slabs to bound float64 temporaries (edge 288: 24M points)
    lin = torch.linspace(-0.5 * _A, 0.5 * _A, edge + 1, dtype=torch.float64, device=device)[:-1]
    keep = []
    slab = max(1, (1 << 22) // (edge * edge))
    for i0 in range(0, edge, slab):
        gi = lin[i0:i0 + slab]
        cu = torch.stack(torch.meshgrid(gi, lin, lin, indexing="ij"), -1).reshape(-1, 3)
        q = cu2qu(cu)
        keep.append(q[ori_in_fz_laue(q, laue_id)])
    out = torch.cat(keep, 0)
    if permute:
        perm = torch.randperm(out.shape[0], generator=generator, device="cpu" if generator is None else generator.device)
        out = out[perm.to(out.device)]
    return out


def sample_ori_fz_laue(laue_id: int, target_n_samples: int, device: torch.device = torch.device("cpu"),
                       permute: bool = True, generator: Optional[torch.Generator] = None) -> torch.Tensor:
    """Cubochoric FZ sampler by target count (laue_fz_ori.py:472-513):
    edge = floor((N * mult / 2)^(1/3)); grid points outside the FZ are
    rejected; optional permutation (seedable through `generator`)."""
    edge = int((target_n_samples * 0.5 * get_laue_mult(laue_id)) ** (1.0 / 3.0))
    return _fz_grid(laue_id, edge, device, permute, generator)


def sample_ori_fz_laue_angle(laue_id: int, angular_resolution_deg: float, device: torch.device = torch.device("cpu"),
                             permute: bool = True, generator: Optional[torch.Generator] = None) -> torch.Tensor:
    """Cubochoric FZ sampler by angular resolution (laue_fz_ori.py:516-557,
    same empirical edge-length fit)."""
    n = (2 * 131.97049 / (angular_resolution_deg - 0.03732) + 1) ** 3
    edge = int((n / (0.5 * get_laue_mult(laue_id))) ** (1.0 / 3.0) + 1.0)
    return _fz_grid(laue_id, edge, device, permute, generator)


def so3_uniform_quat(n: int, device: torch.device = torch.device("cpu"), generator: Optional[torch.Generator] = None,
                     dtype=torch.float32) -> torch.Tensor:
    """Uniform random rotations (Shoemake, Graphics Gems III, 1992), the
    construction of sampling.py:135-176; w may be negative."""
    u, v, w = torch.rand(3, n, generator=generator, dtype=torch.float64).to(device)
    a, b = torch.sqrt(1 - u), torch.sqrt(u)
    q = torch.stack([a * torch.sin(2 * math.pi * v), a * torch.cos(2 * math.pi * v),
                     b * torch.sin(2 * math.pi * w), b * torch.cos(2 * math.pi * w)], 1)
    return q.to(dtype)


def disori_angle_laue(q1: torch.Tensor, q2: torch.Tensor, laue_id: int) -> torch.Tensor:
    """Disorientation angle (radians) between orientations of the same Laue
    group (laue_fz_ori.py:414-469 semantics)."""
    ops = laue_elements(laue_id).to(device=q1.device, dtype=q1.dtype)
    mis = qu_prod(q1, qu_conj(q2))
    # real part of g_a * mis * g_b for all operator pairs, batched
    left = qu_prod(ops[None, :, None, :], mis[:, None, None, :])
    real = qu_prod(left, ops[None, None, :, :])[..., 0].abs()
    c = real.reshape(real.shape[0], -1).max(-1).values
    return 2.0 * torch.acos(c.clamp(-1.0, 1.0))


def ori_to_fz_laue(q: torch.Tensor, laue_id: int, left: bool = False) -> torch.Tensor:
    """Move orientations into the fundamental zone (laue_fz_ori.py:35-81):
    choose the symmetry-equivalent q * g with the largest |w|. With
    left=True the equivalents are g * q instead: the same zone (the real part
    of g * q equals that of q * g) but the representative whose detector
    pattern equals q's, since a pattern is unchanged when its orientation is
    multiplied by a crystal symmetry on the left."""
    ops = laue_elements(laue_id).to(device=q.device, dtype=q.dtype)
    eq = qu_prod(ops[None], q[:, None, :]) if left else qu_prod(q[:, None, :], ops[None])
    best = eq[..., 0].abs().argmax(-1)
    return qu_std(eq[torch.arange(q.shape[0], device=q.device), best])


def get_local_orientation_grid(semi_edge_in_degrees: float, kernel_radius_in_steps: int,
                               axial_grid_dimension: int = 3) -> torch.Tensor:
    """Cubochoric coordinates (n, 3) float32 of a cube of (2r + 1)^3 points
    centred on the identity with semi-edge semi_edge_in_degrees * pi^(2/3) /
    360 (refining.py:13-68); axial_grid_dimension 2 keeps the points with at
    least one zero coordinate, 1 those with at least two."""
    a = semi_edge_in_degrees * math.pi ** (2.0 / 3.0) / 360.0
    t = torch.linspace(-a, a, 2 * kernel_radius_in_steps + 1)
    cu = torch.stack(torch.meshgrid(t, t, t, indexing="ij"), -1).reshape(-1, 3)
    if axial_grid_dimension == 3:
        return cu
    zeros = (cu == 0).sum(-1)
    if axial_grid_dimension == 2:
        return cu[zeros >= 1]
    if axial_grid_dimension == 1:
        return cu[zeros >= 2]
    raise ValueError("axial_grid_dimension must be 1, 2, or 3.")


def local_orientation_grid_quats(semi_edge_in_degrees: float, kernel_radius_in_steps: int,
                                 axial_grid_dimension: int = 3) -> torch.Tensor:
    """The local grid as unit quaternions (n, 4) float32 (cu2qu, w >= 0)."""
    cu = get_local_orientation_grid(semi_edge_in_degrees, kernel_radius_in_steps, axial_grid_dimension)
    return cu2qu(cu.double())


def local_ball_grid_quats(radius_deg: float = 3.0, radius_steps: int = 6) -> torch.Tensor:
    """The local cubochoric grid (get_local_orientation_grid, (2r + 1)^3
    points) as unit quaternions (n, 4) float32, with the semi-edge chosen so
    that its outermost shell lies at exactly radius_deg. The cubochoric map
    sends concentric cubes to concentric balls, so the grid is an
    equal-volume sampling of the ball of rotations within radius_deg of the
    identity (r = 6: 2,197 orientations, about 0.37 degree apart for 3
    degrees)."""
    target = math.radians(radius_deg)
    semi = radius_deg
    for _ in range(6):  # the shell angle is nearly linear in the semi-edge
        a = semi * math.pi ** (2.0 / 3.0) / 360.0
        q = cu2qu(torch.tensor([[a, 0.0, 0.0]], dtype=torch.float64))
        ang = 2.0 * math.atan2(float(q[0, 1:].double().norm()), float(q[0, 0]))
        semi *= target / ang
    return cu2qu(get_local_orientation_grid(semi, radius_steps).double()).contiguous()
