"""Gradient refinement on the GPU (refining.py:247-388): the direction
gradient of the master-pattern lookup (ebsd_sample_dirs_grad) against the
oracle's autograd, and orientation_grad_refinement against the reference's
own run and the oracle."""

import numpy as np
import pytest
import torch

from oracle import di_oracle as O

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _smooth_points(n, seed, mp_shape, align):
    """Unit directions away from the seams of the Lambert map (|x| = |y|,
    z = 0) and, for bilinear, from texel boundaries, where the derivative
    jumps and one float32 ulp of the coordinate may pick either side."""
    g = torch.Generator().manual_seed(seed)
    p = torch.randn(n, 3, generator=g)
    p = p / p.norm(dim=-1, keepdim=True)
    keep = (p[:, 2].abs() > 1e-3) & ((p[:, 0].abs() - p[:, 1].abs()).abs() > 1e-3)
    sq = O.rosca_lambert_side_by_side(p).double()
    H, W = mp_shape
    for c, size in ((sq[:, 0], W), (sq[:, 1], H)):
        x = (c + 1) / 2 * (size - 1) if align else ((c + 1) * size - 1) / 2
        frac = x - torch.floor(x)
        keep &= (frac > 1e-3) & (frac < 1 - 1e-3)
    return p[keep]


@pytest.mark.parametrize("mode,pad,align", [("bicubic", "border", True), ("bicubic", "zeros", False),
                                            ("bicubic", "reflection", True), ("bilinear", "border", False),
                                            ("bilinear", "border", True), ("bilinear", "reflection", False),
                                            ("bilinear", "zeros", True), ("nearest", "border", False)])
def test_lookup_gradient_matches_oracle_autograd(mode, pad, align):
    from ebsdtorch_amd.master_pattern import MasterPattern
    from ebsdtorch_amd.synthetic import band_master_pattern

    mp = band_master_pattern(101, laue_id=11, n_families=8, sigma_deg=2.0, seed=3)
    pts = _smooth_points(40_000, 9, mp.shape, align)
    w = torch.randn(len(pts), generator=torch.Generator().manual_seed(4))  # a generic upstream gradient
    print("lookup gradient", mode, pad, align, len(pts), flush=True)
    d = pts.to(DEV).requires_grad_(True)
    val = MasterPattern(mp.to(DEV), laue_group=11).interpolate(d, mode=mode, padding_mode=pad, align_corners=align,
                                                               normalize_coords=False)
    (val * w.to(DEV)).sum().backward()
    r = pts.clone().requires_grad_(True)
    ref = O.interpolate(mp, r, normalize_coords=False, mode=mode, padding_mode=pad, align_corners=align)
    (ref * w).sum().backward()
    torch.testing.assert_close(val.detach().cpu(), ref.detach(), atol=4e-5, rtol=0)
    scale = r.grad.abs().max()
    if mode == "nearest":
        assert torch.count_nonzero(d.grad) == 0 and torch.count_nonzero(r.grad) == 0
    else:
        assert scale > 0
        torch.testing.assert_close(d.grad.cpu(), r.grad, atol=2e-4 * float(scale), rtol=2e-3)


def _rays8():
    from ebsdtorch_amd.geometry import EBSDGeometry

    return EBSDGeometry((8, 8), proj_center=(0.5, 0.5, 0.6))


@pytest.mark.parametrize("tag,md", [("f32", torch.float32), ("f16", torch.float16)])
def test_grad_refinement_matches_reference_run(golden, tag, md):
    """The reference's fixture (16 patterns, the last four anti-matches): the
    matched patterns keep their orientations exactly, the anti-matched ones
    move as the reference moved them."""
    from ebsdtorch_amd.experiment_pats import ExperimentPatterns
    from ebsdtorch_amd.master_pattern import MasterPattern
    from ebsdtorch_amd.refining import orientation_grad_refinement

    start = torch.tensor(golden["refine_start"])
    ep = ExperimentPatterns(torch.tensor(golden["grad_pats"]).to(DEV))
    ep.set_orientations(start.clone().to(DEV))
    ep.phases = torch.zeros(16, dtype=torch.uint8, device=DEV)
    print("grad refinement fixture", tag, flush=True)
    orientation_grad_refinement(MasterPattern(torch.tensor(golden["mp33"]).to(DEV), laue_group=11), _rays8(), ep,
                                batch_size=8, n_iter=10, match_dtype=md)
    got = ep.get_orientations().cpu()
    ref = torch.tensor(golden[f"grad_{tag}_ori"])
    assert torch.equal(got[:12], start[:12])
    # the same rotation up to the sign of the quaternion
    err = 1 - (got[12:] * ref[12:]).sum(-1).abs()
    assert float(err.max()) < 1e-6, err


def test_grad_refinement_matches_oracle_60x60():
    """60 x 60 patterns of a 201 x 201 band pattern, 48 matched (1 degree
    off) and 16 anti-matched, 20 Adam steps, FP32 dots: the same patterns
    are written back, to the oracle's orientations."""
    from ebsdtorch_amd.experiment_pats import ExperimentPatterns
    from ebsdtorch_amd.geometry import EBSDGeometry
    from ebsdtorch_amd.master_pattern import MasterPattern
    from ebsdtorch_amd.refining import orientation_grad_refinement
    from ebsdtorch_amd.synthetic import band_master_pattern

    mp = band_master_pattern(201, laue_id=11, n_families=12, sigma_deg=2.0, seed=6)
    geom = EBSDGeometry((60, 60), proj_center=(0.5, 0.5, 0.6))
    rays = geom.get_coords_sample_frame((1, 1)).view(-1, 3)
    rays = rays / rays.norm(dim=-1, keepdim=True)
    g = torch.Generator().manual_seed(8)
    truth = O.ori_to_fz_laue(torch.nn.functional.normalize(torch.randn(64, 4, generator=g), dim=1), 11)
    pats = O.interpolate(mp, O.qu_apply(truth[:, None], rays[None]), normalize_coords=False, mode="bicubic",
                         align_corners=True).view(64, 60, 60)
    pats[48:] = pats[48:].max() - pats[48:]
    axis = torch.nn.functional.normalize(torch.randn(64, 3, generator=g), dim=1)
    ang = torch.deg2rad(torch.tensor(1.0))
    dq = torch.cat([torch.cos(ang / 2).expand(64, 1), axis * torch.sin(ang / 2)], 1)
    start = O.qu_prod(dq, truth)
    start = start * torch.where(start[:, :1] < 0, -1.0, 1.0)
    ref = O.orientation_grad_refinement(mp, 11, rays, pats, start, batch_size=32, n_iter=20,
                                        match_dtype=torch.float32)
    ep = ExperimentPatterns(pats.to(DEV))
    ep.set_orientations(start.clone().to(DEV))
    ep.phases = torch.zeros(64, dtype=torch.uint8, device=DEV)
    print("grad refinement 60x60", flush=True)
    orientation_grad_refinement(MasterPattern(mp.to(DEV), laue_group=11), geom, ep, batch_size=32, n_iter=20,
                                match_dtype=torch.float32)
    got = ep.get_orientations().cpu()
    moved_ref = (ref - start).abs().amax(1) > 0
    moved_got = (got - start).abs().amax(1) > 0
    assert torch.equal(moved_ref, moved_got) and int(moved_ref.sum()) == 16
    err = 1 - (got[moved_ref] * ref[moved_ref]).sum(-1).abs()
    assert float(err.max()) < 1e-6, err
