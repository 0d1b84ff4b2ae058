"""Orientation refinement on a local cubochoric grid (ebsd/refining.py).

`orientation_grid_refinement` mirrors refining.py:71-244: per pattern, n_iter
iterations of "compose the local grid with the current orientation, move the
candidates to the fundamental zone, project them, take the best zero-mean
dot product, recentre on it, and shrink the grid where the centre won". Each
iteration is one launch of ebsd_grid_refine_step over a batch: candidate
patterns are generated and scored on the fly, never stored. The per-pattern
grid shrink is a per-pattern level into a small table of the grid at each
shrink level (the grid after k shrinks, k < n_iter), mapped to quaternions
once per call.
"""

from __future__ import annotations

from typing import List, Union

import torch

from . import kernels
from ._lib import DTYPE_F32
from .experiment_pats import ExperimentPatterns
from .geometry import EBSDGeometry
from .master_pattern import MasterPattern
from .orientations import cu2qu, get_local_orientation_grid, laue_elements, ori_to_fz_laue, qu_apply


def _grid_levels(semi_edge_in_degrees: float, kernel_radius_in_steps: int, axial_grid_dimension: int,
                 shrink_factor: float, n_levels: int) -> torch.Tensor:
    """(n_levels, G, 4) float32: cu2qu of the cubochoric grid multiplied by
    shrink_factor k times (float32 multiplications, as the in-place *= of
    refining.py:234-237), k = 0 .. n_levels - 1."""
    cu = get_local_orientation_grid(semi_edge_in_degrees, kernel_radius_in_steps, axial_grid_dimension)
    out = []
    cur = cu.clone()
    for _ in range(n_levels):
        out.append(cu2qu(cur))
        cur = cur * torch.tensor(shrink_factor, dtype=torch.float32)
    return torch.stack(out, 0).contiguous()


def orientation_grid_refinement(
    master_patterns: Union[MasterPattern, List[MasterPattern]],
    geometry: EBSDGeometry,
    experiment_patterns: ExperimentPatterns,
    grid_semi_edge_in_degrees: float,
    batch_size: int,
    virtual_binning: int = 1,
    n_iter: int = 3,
    axial_grid_dimension: int = 3,
    kernel_radius_in_steps: int = 1,
    shrink_factor: float = 0.5,
    average_pattern_center: bool = True,
    match_dtype: torch.dtype = torch.float16,
) -> List[torch.Tensor]:
    """Refine the orientations of indexed patterns in place
    (refining.py:71-244); returns, per phase, the best dot product of every
    pattern at every iteration (n_patterns, n_iter).

    The patterns are scored at the virtually binned size against the
    average-pooled master pattern with the align_corners=True lookup, as the
    reference does. Dot products are accumulated in FP32 whatever
    `match_dtype` (the reference's default FP16 einsum is not reproduced;
    FP32 is at least as precise). Per-pattern pattern centres
    (average_pattern_center=False) are not supported.
    """
    if not average_pattern_center:
        raise NotImplementedError("per-pattern pattern centres are not supported; use average_pattern_center=True")
    if n_iter < 1 or batch_size < 1 or virtual_binning < 1:
        raise ValueError("n_iter, batch_size and virtual_binning must be positive")
    if not hasattr(experiment_patterns, "orientations") or experiment_patterns.orientations is None:
        raise ValueError("The experiment patterns must be indexed before refinement.")
    if isinstance(master_patterns, MasterPattern):
        master_patterns = [master_patterns]
    dev = experiment_patterns.patterns.device
    if dev.type != "cuda":
        raise ValueError("grid refinement runs on the GPU: the experiment patterns must be on a CUDA device")
    rays = geometry.get_coords_sample_frame(binning=(virtual_binning, virtual_binning)).view(-1, 3)
    rays = (rays / rays.norm(dim=-1, keepdim=True)).to(dev, torch.float32).contiguous()
    grids = _grid_levels(grid_semi_edge_in_degrees, kernel_radius_in_steps, axial_grid_dimension, shrink_factor,
                         n_iter).to(dev)
    n_grid = grids.shape[1]
    center = (n_grid - 1) / 2
    all_dots = []
    for phase, indices in enumerate(experiment_patterns.get_indices_per_phase()):
        mp = master_patterns[phase]
        quads = kernels.mp_quad_table(mp.pattern_for_binning(virtual_binning).to(dev), torch.float32)
        ops = laue_elements(mp.laue_group).to(dev, torch.float32).contiguous()
        indices = indices.reshape(-1).to(dev)
        dots = torch.zeros((experiment_patterns.n_patterns, n_iter), dtype=torch.float32, device=dev)
        for batch in torch.split(indices, batch_size):
            pats = experiment_patterns.get_patterns(batch, binning=virtual_binning)
            pats = pats.reshape(pats.shape[0], -1).to(torch.float32)
            x = kernels.prep_patterns(pats, DTYPE_F32, zero_mean=True, unit_norm=False)
            q = experiment_patterns.get_orientations(batch).to(dev, torch.float32).contiguous()
            level = torch.zeros(batch.numel(), dtype=torch.int32, device=dev)
            for j in range(n_iter):
                val, idx, q = kernels.grid_refine_step(quads, rays, x, q, grids, level, ops)
                dots[batch, j] = val
                level += (idx == center).to(torch.int32)
            experiment_patterns.set_orientations(q, batch)
        all_dots.append(dots)
    return all_dots


def orientation_grad_refinement(
    master_patterns: Union[MasterPattern, List[MasterPattern]],
    geometry: EBSDGeometry,
    experiment_patterns: ExperimentPatterns,
    batch_size: int,
    virtual_binning: int = 1,
    n_iter: int = 50,
    learning_rate: float = 0.001,
    average_pattern_center: bool = True,
    match_dtype: torch.dtype = torch.float16,
) -> None:
    """Refine orientations by gradient ascent on the cosine (refining.py:247-388),
    in place: per batch, Adam on the quaternions for n_iter steps; the loss
    is minus the mean cosine between each zero-mean unit pattern and its
    bicubic (align_corners=True) simulation, whose derivative by the
    directions comes from ebsd_sample_dirs_grad; the rest (rotation, means,
    norms, the optimizer) is torch on the patterns' device, as in the
    reference. Afterwards the quaternions go to the positive hemisphere and
    the fundamental zone and are written back only where the final dot is
    below minus the first one (refining.py:358-359, 381-388), which a
    matched pattern (positive dots) never satisfies: the reference's rule,
    kept as it is. The dots are taken from `match_dtype` operands with FP32
    accumulation, rounded to `match_dtype`."""
    if not average_pattern_center:
        raise NotImplementedError("per-pattern pattern centres are not supported; use average_pattern_center=True")
    if not hasattr(experiment_patterns, "orientations") or experiment_patterns.orientations is None:
        raise ValueError("The experiment patterns must be indexed before refinement.")
    if isinstance(master_patterns, MasterPattern):
        master_patterns = [master_patterns]
    dev = experiment_patterns.patterns.device
    if dev.type != "cuda":
        raise ValueError("gradient refinement runs on the GPU: the experiment patterns must be on a CUDA device")
    rays = geometry.get_coords_sample_frame(binning=(virtual_binning, virtual_binning)).view(-1, 3)
    rays = (rays / rays.norm(dim=-1, keepdim=True)).to(dev, torch.float32)
    for phase, indices in enumerate(experiment_patterns.get_indices_per_phase()):
        mp = master_patterns[phase]
        for batch in torch.split(indices.reshape(-1).to(dev), batch_size):
            pats = experiment_patterns.get_patterns(batch, binning=virtual_binning)
            pats = pats.reshape(pats.shape[0], -1).to(torch.float32)
            pats = pats - torch.mean(pats, dim=-1, keepdim=True)
            pats = pats / pats.norm(dim=-1, keepdim=True)
            q = experiment_patterns.get_orientations(batch).to(dev, torch.float32).clone().detach()
            q.requires_grad_(True)
            opt = torch.optim.Adam([q], lr=learning_rate)
            initial = None
            for step in range(n_iter):
                rot = qu_apply(q[:, None, :], rays[None, :, :])
                sim = mp.interpolate(rot, mode="bicubic", align_corners=True, normalize_coords=False,
                                     virtual_binning=virtual_binning).reshape(q.shape[0], -1)
                sim = sim - torch.mean(sim, dim=-1, keepdim=True)
                sim = sim / sim.norm(dim=-1, keepdim=True)
                dots = (pats.to(match_dtype).float() * sim.to(match_dtype).float()).sum(-1).to(match_dtype)
                if step == 0:
                    initial = -dots
                loss = -torch.mean(dots)
                opt.zero_grad()
                loss.backward()
                opt.step()
                with torch.no_grad():
                    q /= q.norm(dim=-1, keepdim=True)
            q = q.detach()
            q = q * ((q[..., :1] >= 0).float() * 2 - 1)
            q = ori_to_fz_laue(q, mp.laue_group)
            upd = dots < initial
            experiment_patterns.set_orientations(q[upd], batch[upd])
