"""The oracle of orientation_grad_refinement (refining.py:247-388) against
the reference's own run (fixture written by tests/golden/make_golden.py), and
the derivative it relies on (bicubic grid_sample through the Lambert map)
against finite differences."""

import numpy as np
import pytest
import torch

from oracle import di_oracle as O


def _rays8():
    from ebsdtorch_amd.geometry import EBSDGeometry

    rays = EBSDGeometry((8, 8), proj_center=(0.5, 0.5, 0.6)).get_coords_sample_frame((1, 1)).view(-1, 3)
    return rays / rays.norm(dim=-1, keepdim=True)


@pytest.mark.parametrize("tag,md", [("f32", torch.float32), ("f16", torch.float16)])
def test_grad_refinement_matches_reference(golden, tag, md):
    got = O.orientation_grad_refinement(torch.tensor(golden["mp33"]), 11, _rays8(), torch.tensor(golden["grad_pats"]),
                                        torch.tensor(golden["refine_start"]), batch_size=8, n_iter=10, match_dtype=md)
    np.testing.assert_array_equal(got.numpy(), golden[f"grad_{tag}_ori"])
    # only the anti-matched patterns (inverted contrast) are written back
    changed = np.flatnonzero(np.abs(got.numpy() - golden["refine_start"]).max(1) > 0)
    assert changed.tolist() == [12, 13, 14, 15]


@pytest.mark.parametrize("mode,pad,align", [("bicubic", "border", True), ("bicubic", "zeros", False),
                                            ("bilinear", "border", True), ("bilinear", "reflection", False)])
def test_interpolate_derivative_by_finite_differences(mode, pad, align):
    """torch.autograd.gradcheck of the oracle's interpolate (float64) in the
    directions: the Lambert map's and grid_sample's derivatives."""
    g = torch.Generator().manual_seed(2)
    mp = torch.rand(9, 18, generator=g, dtype=torch.float64)
    pts = torch.randn(40, 3, generator=g, dtype=torch.float64)
    pts = pts / pts.norm(dim=-1, keepdim=True)
    # away from the seams where the map or the sampling is not differentiable
    keep = (pts[:, 2].abs() > 0.05) & ((pts[:, 0].abs() - pts[:, 1].abs()).abs() > 0.05)
    pts = pts[keep].clone().requires_grad_(True)

    def f(p):
        return O.interpolate(mp, p, normalize_coords=False, mode=mode, padding_mode=pad, align_corners=align)

    assert torch.autograd.gradcheck(f, (pts,), eps=1e-7, atol=1e-5, rtol=1e-4)
