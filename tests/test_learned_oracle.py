"""Pin the learned-proposal oracle (CPU): Philox known answers, noise
statistics, 6D rotation and loss closed forms, loss gradient against finite
differences, convolution against a brute-force loop, local scores against the
DI oracle's qu_apply projection. Also the product's host-side pieces that need
no GPU: the local grid, the left fundamental-zone reduction and the
pattern-centre Jacobian."""

import math

import numpy as np
import pytest
import torch

from oracle import di_oracle as O
from oracle import learned_oracle as L


# ------------------------------------------------------------------ Philox


@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox_known_answers(ctr, key, expect):
    got = L.philox4x32_10(np.array([ctr], dtype=np.uint32), key)[0]
    assert tuple(int(v) for v in got) == expect


# ------------------------------------------------------------------ noise


def _counts(lam: float, n: int, seed: int):
    clean = np.ones((1, n), dtype=np.float32)
    counts, _ = L.poisson_gaussian_noise(clean, np.array([lam], np.float32), 0.0, seed)
    return counts[0]


@pytest.mark.parametrize("lam", [0.7, 3.0, 9.5, 10.5, 60.0, 300.0])
def test_poisson_moments(lam):
    n = 3000
    k = _counts(lam, n, seed=11 + int(lam))
    mean_tol = 5 * math.sqrt(lam / n)
    assert abs(k.mean() - lam) < mean_tol
    # variance of a Poisson sample: sd of s^2 is about sqrt((lam + 2 lam^2) / n)
    assert abs(k.var(ddof=1) - lam) < 5 * math.sqrt((lam + 2 * lam * lam) / n)


@pytest.mark.parametrize("lam", [3.0, 25.0])
def test_poisson_chi_square(lam):
    n = 4000
    k = _counts(lam, n, seed=5)
    lo, hi = max(0, int(lam - 3 * math.sqrt(lam))), int(lam + 3 * math.sqrt(lam)) + 1
    pmf = np.array([math.exp(-lam + j * math.log(lam) - math.lgamma(j + 1)) for j in range(lo, hi + 1)])
    obs = np.array([(k == j).sum() for j in range(lo, hi + 1)], dtype=np.float64)
    exp = pmf * n
    chi2 = ((obs - exp) ** 2 / exp).sum()
    dof = len(pmf)
    assert chi2 < dof + 5 * math.sqrt(2 * dof)


def test_gaussian_part_and_scaling():
    n = 3000
    clean = np.full((1, n), 0.25, dtype=np.float32)
    counts, out = L.poisson_gaussian_noise(clean, np.array([1e-9], np.float32), 0.5, seed=3)
    # rate ~ 0 -> counts 0, out = N(0, 0.5^2)
    assert counts.max() == 0
    assert abs(out.mean()) < 5 * 0.5 / math.sqrt(n)
    assert abs(out.std() - 0.5) < 0.05
    # scaling: counts / c averages to p / mean(p) (= 1 here)
    c2, o2 = L.poisson_gaussian_noise(clean, np.array([50.0], np.float32), 0.0, seed=4)
    np.testing.assert_allclose(o2, c2 / 50.0, rtol=1e-6)
    assert abs(o2.mean() - 1.0) < 5 * math.sqrt(50.0 / n) / 50.0


def test_noise_streams_are_counter_based():
    clean = np.ones((2, 16), dtype=np.float32)
    a = L.poisson_gaussian_noise(clean, np.array([5.0, 5.0], np.float32), 0.1, seed=9, offset=0)[1]
    b = L.poisson_gaussian_noise(clean[1:], np.array([5.0], np.float32), 0.1, seed=9, offset=0)[1]
    c = L.poisson_gaussian_noise(clean, np.array([5.0, 5.0], np.float32), 0.1, seed=9, offset=1)[1]
    # row 1 of a batch differs from row 0 of another batch (row is counted)
    assert not np.array_equal(a[1], b[0])
    assert not np.array_equal(a, c)


# ------------------------------------------------------------------ 6D / loss


def _rand_quats(n, seed):
    g = torch.Generator().manual_seed(seed)
    q = torch.randn(n, 4, generator=g, dtype=torch.float64)
    q = q / q.norm(dim=1, keepdim=True)
    return torch.where(q[:, :1] < 0, -q, q)


def test_rot6d_matrix_properties():
    g = torch.Generator().manual_seed(0)
    z = torch.randn(64, 6, generator=g, dtype=torch.float64)
    R = L.rot6d_matrix(z)
    eye = torch.eye(3, dtype=torch.float64).expand(64, 3, 3)
    torch.testing.assert_close(R @ R.transpose(1, 2), eye, atol=1e-10, rtol=0)
    torch.testing.assert_close(torch.linalg.det(R), torch.ones(64, dtype=torch.float64), atol=1e-10, rtol=0)
    # the first two rows of a rotation map back to it
    Q = O.qu2om(_rand_quats(64, 1))
    torch.testing.assert_close(L.rot6d_matrix(Q[:, :2].reshape(64, 6)), Q, atol=1e-10, rtol=0)
    # identity
    torch.testing.assert_close(L.rot6d_matrix(torch.tensor([[2.0, 0, 0, 0.5, 3.0, 0]], dtype=torch.float64))[0],
                               torch.eye(3, dtype=torch.float64))


def _axis_angle_matrix(axis, theta):
    a = torch.as_tensor(axis, dtype=torch.float64)
    a = a / a.norm()
    q = torch.cat([torch.tensor([math.cos(theta / 2)], dtype=torch.float64), a * math.sin(theta / 2)])
    return O.qu2om(q[None])[0]


@pytest.mark.parametrize("theta_deg", [0.5, 5.0, 17.0, 29.0])
def test_geodesic_loss_closed_form(theta_deg):
    ops = O.laue_elements(11).double()
    qt = _rand_quats(8, 2)
    Rt = O.qu2om(qt)
    th = math.radians(theta_deg)
    Rp = torch.stack([_axis_angle_matrix([1.0, 2.0, 0.5 + i], th) @ Rt[i] for i in range(8)])
    z = Rp[:, :2].reshape(8, 6)
    loss = L.geodesic_sym_loss(z, qt, ops)
    torch.testing.assert_close(loss, torch.full((8,), th, dtype=torch.float64), atol=2e-6, rtol=0)


def test_geodesic_loss_symmetry_invariance():
    ops = O.laue_elements(11).double()
    qt = _rand_quats(24, 3)
    # predict G * R_true for every operator G: loss hits the clamp floor
    G = O.qu2om(ops)
    Rt = O.qu2om(qt)
    z = (G @ Rt)[:, :2].reshape(24, 6)
    loss = L.geodesic_sym_loss(z, qt, ops)
    floor = math.acos(1 - 1e-6)
    torch.testing.assert_close(loss, torch.full((24,), floor, dtype=torch.float64), atol=1e-7, rtol=0)
    # right-multiplying by a symmetry is not a symmetry of the pattern: no floor
    zr = (Rt @ G[5])[:, :2].reshape(24, 6)
    assert (L.geodesic_sym_loss(zr, qt, ops) > 1e-2).sum() > 0


def test_geodesic_loss_gradient_finite_differences():
    ops = O.laue_elements(11).double()
    qt = _rand_quats(6, 4)
    g = torch.Generator().manual_seed(5)
    z = torch.randn(6, 6, generator=g, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: L.geodesic_sym_loss(t, qt, ops).sum(), (z,), eps=1e-6, atol=1e-6)


# ------------------------------------------------------------------ network


def test_conv_matches_bruteforce():
    g = torch.Generator().manual_seed(0)
    x = torch.randn(2, 3, 7, 6, generator=g, dtype=torch.float64)
    w = torch.randn(4, 3, 3, 3, generator=g, dtype=torch.float64)
    b = torch.randn(4, generator=g, dtype=torch.float64)
    ref = L.conv_bruteforce(x, w, b)
    got = torch.nn.functional.conv2d(x, w, b, stride=2, padding=1)
    torch.testing.assert_close(got, ref, atol=1e-12, rtol=0)


def test_net_forward_shapes_and_flatten_order():
    g = torch.Generator().manual_seed(1)
    H = W = 16
    chans = [4, 6, 8, 8]
    params, ci = {}, 1
    for l, co in enumerate(chans):
        params[f"conv{l}.w"] = torch.randn(co, ci, 3, 3, generator=g, dtype=torch.float64) * 0.3
        params[f"conv{l}.b"] = torch.randn(co, generator=g, dtype=torch.float64) * 0.1
        ci = co
    feat = 1 * 1 * 8
    params["fc1.w"] = torch.randn(16, feat, generator=g, dtype=torch.float64)
    params["fc1.b"] = torch.zeros(16, dtype=torch.float64)
    params["fc2.w"] = torch.randn(6, 16, generator=g, dtype=torch.float64)
    params["fc2.b"] = torch.zeros(6, dtype=torch.float64)
    x = torch.randn(3, H * W, generator=g, dtype=torch.float64)
    z = L.net_forward(params, x, (H, W))
    assert z.shape == (3, 6)
    # NHWC flatten at 2 x 2: feature index (y * 2 + x) * C + c
    H = W = 32
    x = torch.randn(2, H * W, generator=g, dtype=torch.float64)
    params["fc1.w"] = torch.zeros(16, 2 * 2 * 8, dtype=torch.float64)
    params["fc1.w"][0, (1 * 2 + 0) * 8 + 3] = 1.0  # picks y=1, x=0, c=3
    params["fc2.w"] = torch.zeros(6, 16, dtype=torch.float64)
    params["fc2.w"][0, 0] = 1.0
    h = x.view(-1, 1, H, W) * math.sqrt(H * W)
    for l in range(4):
        h = torch.relu(torch.nn.functional.conv2d(h, params[f"conv{l}.w"], params[f"conv{l}.b"], stride=2,
                                                  padding=1))
    z = L.net_forward(params, x, (H, W))
    torch.testing.assert_close(z[:, 0], torch.relu(h[:, 3, 1, 0]), atol=1e-12, rtol=0)


# ------------------------------------------------------------------ local scores


def test_local_scores_match_qu_apply_projection():
    torch.manual_seed(0)
    S = 41
    mp = torch.rand(S, 2 * S, dtype=torch.float32)
    rays = O.normalized_rays((12, 12), (0.5, 0.5, 0.6))
    centers = _rand_quats(3, 6).float()
    grid = _rand_quats(5, 7).float()
    grid = torch.nn.functional.normalize(grid + torch.tensor([8.0, 0, 0, 0]), dim=1)  # near identity
    x = torch.randn(3, rays.shape[0], dtype=torch.float64)
    x = x - x.mean(1, keepdim=True)
    x = x / x.norm(dim=1, keepdim=True)
    got = L.local_scores(mp, rays, x, centers, grid)
    for i in range(3):
        cand = O.qu_prod(grid.double(), centers[i].double()[None].expand(5, 4))
        cand = cand / cand.norm(dim=1, keepdim=True)
        pats = O.project_dictionary(mp.double(), rays.double(), cand)
        ref = (pats @ x[i]) / pats.norm(dim=1)
        torch.testing.assert_close(got[i], ref, atol=1e-10, rtol=0)


def test_topk_total_order():
    s = torch.tensor([[0.5, 0.9, 0.9, 0.1]], dtype=torch.float64)
    v, i = L.topk_total_order(s, 3)
    assert i.tolist() == [[1, 2, 0]]


# ------------------------------------------------------------------ host-side product pieces


def test_local_ball_grid():
    from ebsdtorch_amd.orientations import local_ball_grid_quats

    g = local_ball_grid_quats(3.0, 6).double()
    assert g.shape == (2197, 4)
    ang = torch.rad2deg(2 * torch.atan2(g[:, 1:].norm(dim=1), g[:, 0]))
    assert abs(float(ang.max()) - 3.0) < 1e-3
    assert float(ang.min()) == 0.0
    # concentric shells: every point on the outer cube shell is at the radius
    r = 6
    t = torch.arange(2 * r + 1) - r
    ijk = torch.stack(torch.meshgrid(t, t, t, indexing="ij"), -1).reshape(-1, 3)
    shell = ijk.abs().amax(-1) == r
    assert torch.allclose(ang[shell], torch.full_like(ang[shell], 3.0), atol=1e-3)
    assert float(ang[~shell].max()) < 3.0 * (r - 1) / r + 1e-3


def test_left_fz_reduction_keeps_patterns():
    from ebsdtorch_amd.orientations import ori_in_fz_laue, ori_to_fz_laue

    q = _rand_quats(32, 8).float()
    ql = ori_to_fz_laue(q, 11, left=True)
    qr = ori_to_fz_laue(q, 11)
    assert bool(ori_in_fz_laue(ql, 11).all()) and bool(ori_in_fz_laue(qr, 11).all())
    # the left representative projects to the same pattern on a symmetric master pattern
    from ebsdtorch_amd.synthetic import band_master_pattern

    # (bilinear sampling of the discretised Lambert image is symmetric only up
    # to its resolution: ~1e-2 here, against O(1) for a non-symmetry)
    mp = band_master_pattern(201, laue_id=11, n_families=6, sigma_deg=5.0, seed=1)
    rays = O.normalized_rays((10, 10), (0.5, 0.5, 0.6))
    a = O.project_dictionary(mp, rays, q)
    b = O.project_dictionary(mp, rays, ql)
    c = O.project_dictionary(mp, rays, qr)
    assert float((a - b).abs().max()) < 2e-2
    assert float((a - c).abs().max()) > 0.3


def test_pc_jacobian_matches_geometry():
    from ebsdtorch_amd.geometry import EBSDGeometry

    g0 = EBSDGeometry((20, 24), proj_center=(0.5, 0.5, 0.6))
    J = g0.pc_jacobian().double()
    d = torch.tensor([0.01, -0.007, 0.004], dtype=torch.float64)
    g1 = EBSDGeometry((20, 24), proj_center=(0.51, 0.493, 0.604))
    c0 = g0.get_coords_sample_frame((1, 1)).double()
    c1 = g1.get_coords_sample_frame((1, 1)).double()
    torch.testing.assert_close(c1 - c0, (J @ d).expand_as(c0), atol=2e-5, rtol=0)
