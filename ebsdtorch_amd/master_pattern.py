"""Master pattern container (ebsd/master_pattern.py:22-209).

The pattern is the side-by-side Rosca-Lambert image (S, 2S): northern
hemisphere in columns [0, S), southern in [S, 2S). interpolate() runs the
HIP lookup kernel (ebsd_sample_dirs); the dictionary path does not call it but
fuses the same lookup into ebsd_project.
"""

from __future__ import annotations

from typing import Optional, Tuple, Union

import torch

from . import _lib
from ._lib import check, require_device


def space_group_to_laue(space_group: int) -> int:
    """Space group number -> Laue group id (utils/symmetry_classes.py:4-30),
    the reference's table as it is: it maps 149-155 (point group 32) to Laue
    group 6 and 194 (P6_3/mmc) to 10, where the International Tables have 7
    and 9."""
    if space_group < 1 or space_group > 230:
        raise ValueError("The space group must be between 1 and 230 inclusive.")
    for bound, lid in ((206, 11), (193, 10), (176, 9), (167, 8), (155, 7), (142, 6), (88, 5), (74, 4), (15, 3),
                       (2, 2)):
        if space_group > bound:
            return lid
    return 1


_POINT_GROUPS = {
    11: ("m3-m", "4-3m", "432"), 10: ("m3-", "23"), 9: ("6/mmm", "6-m2", "6mm", "622"), 8: ("6/m", "6-", "6"),
    7: ("3-m", "3m", "32"), 6: ("3-", "3"), 5: ("4/mmm", "4-2m", "4mm", "422"), 4: ("4/m", "4-", "4"),
    3: ("mmm", "mm2", "222"), 2: ("2/m", "m", "2"), 1: ("1-", "1"),
}


def point_group_to_laue(point_group: str) -> int:
    """Hermann-Mauguin point group -> Laue group id (symmetry_classes.py:33-64)."""
    for lid, names in _POINT_GROUPS.items():
        if point_group in names:
            return lid
    raise ValueError(f"The point group symbol {point_group!r} is not recognized")


class MasterPattern(torch.nn.Module):
    def __init__(self, master_pattern: Union[torch.Tensor, Tuple[torch.Tensor, torch.Tensor]],
                 laue_group: Optional[int] = None, point_group: Optional[str] = None,
                 space_group: Optional[int] = None):
        super().__init__()
        given = [v is not None for v in (laue_group, point_group, space_group)]
        if sum(given) == 0:
            raise ValueError("At least one of laue_groups, point_groups, or space_groups must be provided")
        if sum(given) > 1:
            raise ValueError("Only one of laue_groups, point_groups, or space_groups should be provided")
        if isinstance(master_pattern, tuple):
            if master_pattern[0].dim() != 2 or master_pattern[1].dim() != 2:
                raise ValueError("Both tensors in 'master_pattern' must be 2D tensors when provided as a tuple")
            master_pattern = torch.cat(master_pattern, dim=-1)
        if master_pattern.dim() != 2:
            raise ValueError(f"'master_pattern' must be a 2D tensor, got shape {tuple(master_pattern.shape)}")
        self.register_buffer("master_pattern", master_pattern)
        if laue_group is not None:
            self.laue_group = int(laue_group)
        elif point_group is not None:
            self.laue_group = point_group_to_laue(point_group)
        else:
            self.laue_group = space_group_to_laue(space_group)
        self.master_pattern_binned = None

    def bin(self, binning_factor: int) -> None:
        """Average-pool the master pattern (master_pattern.py:94-113)."""
        self.master_pattern_binned = torch.nn.functional.avg_pool2d(
            self.master_pattern[None, None], binning_factor, stride=binning_factor)[0, 0]

    def normalize(self, norm_type: str) -> None:
        """"minmax", "zeromean" or "standard", in place (master_pattern.py:115-139)."""
        mp = self.master_pattern
        if norm_type == "minmax":
            lo, hi = torch.min(mp), torch.max(mp)
            mp -= lo
            mp /= 1e-4 + hi - lo
        elif norm_type == "zeromean":
            mp -= torch.mean(mp)
        elif norm_type == "standard":
            mp -= torch.mean(mp)
            mp /= torch.std(mp)
        else:
            raise ValueError(f"Invalid normalization method. Got {norm_type} but expected 'minmax', "
                             "'zeromean', or 'standard'.")

    def pattern_for_binning(self, virtual_binning: int) -> torch.Tensor:
        if virtual_binning > 1:
            if self.master_pattern_binned is None:
                self.bin(virtual_binning)
            mp = self.master_pattern_binned
        else:
            mp = self.master_pattern
        return mp.contiguous().float()

    def interpolate(self, coords_3D: torch.Tensor, mode: str = "bilinear", padding_mode: str = "border",
                    align_corners: bool = False, normalize_coords: bool = True,
                    virtual_binning: int = 1) -> torch.Tensor:
        """Master-pattern values at sphere points (..., 3) -> (...)
        (master_pattern.py:141-209): the side-by-side Lambert coordinates of
        each point sampled as grid_sample samples with this mode ("bilinear",
        "nearest", "bicubic"), padding ("zeros", "border", "reflection") and
        corner alignment. The defaults are the dictionary-indexing lookup
        (ebsd_sample_dirs); the other settings run ebsd_sample_dirs_mode."""
        if mode not in _lib.SAMPLE_MODES:
            raise ValueError(f"nn.functional.grid_sample(): expected mode to be 'bilinear', 'nearest' or "
                             f"'bicubic', but got: '{mode}'")
        if padding_mode not in _lib.PADDING_MODES:
            raise ValueError(f"nn.functional.grid_sample(): expected padding_mode to be 'zeros', 'border', or "
                             f"'reflection', but got: '{padding_mode}'")
        require_device(coords_3D, "coords_3D")
        mp = self.pattern_for_binning(virtual_binning).to(coords_3D.device)
        if coords_3D.requires_grad and torch.is_grad_enabled():
            # differentiable in the directions, as grid_sample is in its grid
            c = coords_3D / torch.norm(coords_3D, dim=-1, keepdim=True) if normalize_coords else coords_3D
            out = _SampleDirs.apply(c.reshape(-1, 3).float().contiguous(), mp, _lib.SAMPLE_MODES[mode],
                                    _lib.PADDING_MODES[padding_mode], bool(align_corners))
            return out.view(coords_3D.shape[:-1])
        dirs = coords_3D.reshape(-1, 3).float().contiguous()
        out = torch.empty(dirs.shape[0], dtype=torch.float32, device=dirs.device)
        lib = _lib.load()
        renorm = 1 if normalize_coords else 0
        stream = _lib.stream_ptr(dirs.device)
        if mode == "bilinear" and padding_mode == "border" and not align_corners:
            st = lib.ebsd_sample_dirs(mp.data_ptr(), mp.shape[0], mp.shape[1], dirs.data_ptr(), dirs.shape[0],
                                      out.data_ptr(), renorm, stream)
            check(st, "ebsd_sample_dirs")
        else:
            st = lib.ebsd_sample_dirs_mode(mp.data_ptr(), mp.shape[0], mp.shape[1], dirs.data_ptr(),
                                           dirs.shape[0], out.data_ptr(), renorm, _lib.SAMPLE_MODES[mode],
                                           _lib.PADDING_MODES[padding_mode], 1 if align_corners else 0, stream)
            check(st, "ebsd_sample_dirs_mode")
        return out.view(coords_3D.shape[:-1])


class _SampleDirs(torch.autograd.Function):
    """Master-pattern lookup (n, 3) -> (n,) with its gradient by the
    directions from ebsd_sample_dirs_grad (the derivative autograd takes
    through the Lambert map and grid_sample; none for the master pattern,
    which the reference keeps as a buffer)."""

    @staticmethod
    def forward(ctx, dirs, mp, mode, padding, align):
        out = torch.empty(dirs.shape[0], dtype=torch.float32, device=dirs.device)
        grad = torch.empty_like(dirs)
        st = _lib.load().ebsd_sample_dirs_grad(mp.data_ptr(), mp.shape[0], mp.shape[1], dirs.data_ptr(),
                                               dirs.shape[0], out.data_ptr(), grad.data_ptr(), mode, padding,
                                               1 if align else 0, _lib.stream_ptr(dirs.device))
        check(st, "ebsd_sample_dirs_grad")
        ctx.save_for_backward(grad)
        return out

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return g[:, None] * grad, None, None, None, None
