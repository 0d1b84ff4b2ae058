"""MasterPattern.interpolate in every grid_sample mode on the GPU
(ebsd_sample_dirs_mode) against the reference-written fixtures and the
oracle (master_pattern.py:141-209)."""

import numpy as np
import pytest
import torch

from oracle import di_oracle as O

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)
MODES = [(m, p, a) for m in ("bilinear", "nearest", "bicubic") for p in ("zeros", "border", "reflection")
         for a in (False, True)]
# The Lambert coordinates differ from the oracle's by the float32 rounding of
# the arctangent (<= 1.5 ulp); values follow within this bound on patterns in
# [0, 1.2] (bicubic's weights reach 1.4 in sum of magnitudes).
TOL = {"bilinear": 2e-5, "nearest": 0.0, "bicubic": 4e-5}


def _near_half(mp_shape, pts, align, tol=1e-4):
    """Points whose source coordinate (oracle arithmetic) lies within `tol`
    pixels of a rounding boundary of nearest mode, where one float32 ulp of
    the coordinate may pick the other texel."""
    pts = pts / torch.norm(pts, dim=-1, keepdim=True)
    g = O.rosca_lambert_side_by_side(pts).double()
    H, W = mp_shape
    out = torch.zeros(len(pts), dtype=torch.bool)
    for c, size in ((g[:, 0], W), (g[:, 1], H)):
        x = (c + 1) / 2 * (size - 1) if align else ((c + 1) * size - 1) / 2
        out |= (x - torch.floor(x) - 0.5).abs() < tol
    return out


def _check(got, ref, mode, exempt=None):
    if mode == "nearest":
        bad = got != ref
        if exempt is not None:
            bad &= ~exempt
        assert int(bad.sum()) == 0, (int(bad.sum()), torch.nonzero(bad)[:5].flatten().tolist())
    else:
        torch.testing.assert_close(got, ref, atol=TOL[mode], rtol=0)


@pytest.mark.parametrize("mode,pad,align", MODES)
def test_interpolate_modes_match_reference_fixture(golden, mode, pad, align):
    from ebsdtorch_amd.master_pattern import MasterPattern

    mp = torch.tensor(golden["interp_modes_mp"])
    pts = torch.tensor(golden["interp_modes_pts"])
    print("interpolate fixture", mode, pad, align, flush=True)
    got = MasterPattern(mp.to(DEV), laue_group=11).interpolate(pts.to(DEV), mode=mode, padding_mode=pad,
                                                               align_corners=align, normalize_coords=True).cpu()
    ref = torch.tensor(golden[f"interp_{mode}_{pad}_{int(align)}"])
    _check(got, ref, mode, _near_half(mp.shape, pts, align))


@pytest.mark.parametrize("mode,pad,align", MODES)
def test_interpolate_modes_match_oracle(mode, pad, align):
    """A 101 x 101 band pattern at 50,000 directions, unnormalised and with
    virtual binning 2 (the binned pattern is sampled)."""
    from ebsdtorch_amd.master_pattern import MasterPattern
    from ebsdtorch_amd.synthetic import band_master_pattern

    mp = band_master_pattern(101, laue_id=11, n_families=8, sigma_deg=2.0, seed=3)
    g = torch.Generator().manual_seed(5)
    pts = torch.randn(50_000, 3, generator=g) * (0.5 + torch.rand(50_000, 1, generator=g))
    print("interpolate oracle", mode, pad, align, flush=True)
    m = MasterPattern(mp.to(DEV), laue_group=11)
    got = m.interpolate(pts.to(DEV), mode=mode, padding_mode=pad, align_corners=align).cpu()
    ref = O.interpolate(mp, pts, normalize_coords=True, mode=mode, padding_mode=pad, align_corners=align)
    _check(got, ref, mode, _near_half(mp.shape, pts, align))
    # virtual binning samples the 2 x 2 average-pooled pattern
    got2 = m.interpolate(pts.to(DEV), mode=mode, padding_mode=pad, align_corners=align, virtual_binning=2).cpu()
    binned = torch.nn.functional.avg_pool2d(mp[None, None], 2, stride=2)[0, 0]
    ref2 = O.interpolate(binned, pts, normalize_coords=True, mode=mode, padding_mode=pad, align_corners=align)
    _check(got2, ref2, mode, _near_half(binned.shape, pts, align))


def test_interpolate_rejects_unknown_modes():
    from ebsdtorch_amd.master_pattern import MasterPattern

    m = MasterPattern(torch.rand(9, 18, device=DEV), laue_group=11)
    pts = torch.randn(4, 3, device=DEV)
    with pytest.raises(ValueError):
        m.interpolate(pts, mode="area")
    with pytest.raises(ValueError):
        m.interpolate(pts, padding_mode="wrap")


def test_interpolate_with_a_host_master_pattern():
    """A MasterPattern kept on the host (as the reference builds it) samples
    GPU directions: the buffer goes to the directions' device."""
    from ebsdtorch_amd.master_pattern import MasterPattern

    mp = torch.rand(21, 42, generator=torch.Generator().manual_seed(1))
    pts = torch.randn(500, 3, generator=torch.Generator().manual_seed(2))
    host = MasterPattern(mp, laue_group=11)
    dev = MasterPattern(mp.to(DEV), laue_group=11)
    for kw in ({}, {"mode": "bicubic", "align_corners": True}):
        assert torch.equal(host.interpolate(pts.to(DEV), **kw), dev.interpolate(pts.to(DEV), **kw))
