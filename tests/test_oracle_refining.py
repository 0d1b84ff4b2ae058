"""Pin the oracle's restatement of the orientation grid refinement
(refining.py:13-244) on the CPU: the local grid's closed-form layout, the
fundamental-zone reduction against the golden-pinned FZ test and the group,
the align_corners=True lookup against a scalar bilinear loop, and the
refinement loop on a synthetic scene."""

import math

import pytest
import torch

from oracle import di_oracle as O


@pytest.mark.parametrize("r", [1, 2, 6])
def test_local_grid_layout(r):
    semi = 1.7
    a = semi * math.pi ** (2 / 3) / 360.0
    cu = O.get_local_orientation_grid(semi, r)
    n = 2 * r + 1
    assert cu.shape == (n ** 3, 3)
    # the middle point (the index the refinement tests for a shrink) is the
    # identity up to linspace's rounding
    assert float(cu[(n ** 3 - 1) // 2].abs().max()) < 1e-9
    assert math.isclose(float(cu.max()), a, rel_tol=1e-6) and math.isclose(float(cu.min()), -a, rel_tol=1e-6)
    # 'ij' meshgrid order: the last coordinate runs fastest
    assert float(cu[1, 2] - cu[0, 2]) > 0 and float(cu[1, 0] - cu[0, 0]) == 0
    # the biaxial / uniaxial grids keep points with exact zero coordinates,
    # which exist only where linspace's middle value is exactly 0
    if float(cu[(n ** 3 - 1) // 2].abs().max()) == 0.0:
        assert O.get_local_orientation_grid(semi, r, 2).shape[0] == n ** 3 - (n - 1) ** 3
        assert O.get_local_orientation_grid(semi, r, 1).shape[0] == 1 + 3 * (n - 1)
    from ebsdtorch_amd.orientations import get_local_orientation_grid

    assert torch.equal(get_local_orientation_grid(semi, r), cu)


@pytest.mark.parametrize("laue", [11, 9, 5])
def test_ori_to_fz_laue_is_an_equivalent_inside_the_zone(laue):
    g = torch.Generator().manual_seed(laue)
    q = torch.nn.functional.normalize(torch.randn(500, 4, generator=g, dtype=torch.float64), dim=1)
    r = O.ori_to_fz_laue(q, laue)
    ops = O.laue_elements(laue)
    # r = q * g for some operator g of the group
    cand = O.qu_prod(q[:, None, :], ops[None])
    dist = (cand - r[:, None, :]).abs().amax(-1)
    assert float(dist.min(1).values.max()) < 1e-12
    # the largest |w| over the equivalents q * g
    assert bool((r[:, 0].abs() >= cand[..., 0].abs().amax(1) - 1e-12).all())
    # and inside the fundamental zone (ori_in_fz_laue is pinned by golden vectors)
    assert float(O.ori_in_fz_laue(r, laue).double().mean()) > 0.99


def test_aligned_lookup_matches_scalar_bilinear():
    g = torch.Generator().manual_seed(3)
    S = 7
    mp = torch.rand(S, 2 * S, generator=g, dtype=torch.float64)
    v = torch.nn.functional.normalize(torch.randn(300, 3, generator=g, dtype=torch.float64), dim=1)
    got = O.interpolate_aligned(mp, v)
    uv = O.rosca_lambert_side_by_side(v)
    for i in range(v.shape[0]):
        x, y = float(uv[i, 0]), float(uv[i, 1])
        ix = min(max((x + 1) / 2 * (2 * S - 1), 0.0), 2 * S - 1)
        iy = min(max((y + 1) / 2 * (S - 1), 0.0), S - 1)
        x0, y0 = int(math.floor(ix)), int(math.floor(iy))
        x1, y1 = min(x0 + 1, 2 * S - 1), min(y0 + 1, S - 1)
        wx, wy = ix - x0, iy - y0
        ref = ((1 - wy) * ((1 - wx) * mp[y0, x0] + wx * mp[y0, x1])
               + wy * ((1 - wx) * mp[y1, x0] + wx * mp[y1, x1]))
        assert abs(float(got[i]) - float(ref)) < 1e-12


def test_grid_refinement_converges_and_never_loses_score():
    from ebsdtorch_amd.synthetic import band_master_pattern

    mp = band_master_pattern(101, laue_id=11, n_families=10, sigma_deg=2.0, seed=0)
    g = torch.Generator().manual_seed(0)
    truth = O.ori_to_fz_laue(torch.nn.functional.normalize(torch.randn(8, 4, generator=g), dim=1), 11)
    rays = O.normalized_rays((20, 20), (0.5, 0.5, 0.6)).float()
    pats = O.project_dictionary(mp, rays, truth).view(8, 20, 20) + 0.02 * torch.randn(8, 20, 20, generator=g)
    pert = torch.nn.functional.normalize(torch.tensor([[1.0, 0.01, -0.008, 0.012]]).expand(8, 4), dim=1)
    start = O.qu_prod(pert, truth)
    out, dots = O.orientation_grid_refinement(mp, 11, rays, pats, start, 1.5, 4, n_iter=8)
    err0 = torch.rad2deg(O.disori_angle_laue(start.double(), truth.double(), 11))
    err = torch.rad2deg(O.disori_angle_laue(out.double(), truth.double(), 11))
    assert float(err.max()) < 0.5 * float(err0.min())
    assert bool((dots[:, 1:] >= dots[:, :-1] - 1e-5 * float(dots.abs().max())).all())
    # one iteration with a single-point grid at the identity (semi-edge 0)
    # returns its input
    q1, d1 = O.orientation_grid_refinement(mp, 11, rays, pats, out, 0.0, 4, n_iter=1, kernel_radius_in_steps=0)
    assert torch.allclose(O.qu_std(q1), O.qu_std(out), atol=1e-6)
