"""Pin the product's own sampler, symmetry, disorientation, rotation and
geometry functions (ebsdtorch_amd.orientations / .geometry, which run as
torch ops and never touch the oracle) against the reference's golden
vectors in tests/golden/ref_vectors.npz.

The product's sampler evaluates the cubochoric map in float64 and rounds
once, where the reference evaluates a float32 polynomial fit
(orientations.py:419-516), so sampler outputs are compared to an absolute
bound of 6e-7 (5 float32 ulps of 1.0) with the same points, in the same
order, and q / -q treated as one rotation where w = 0; every other function
is compared bit for bit or to the stated bound."""

import numpy as np
import pytest
import torch

import ebsdtorch_amd as E
from ebsdtorch_amd import orientations as P

T = torch.tensor
SAMPLER_TOL = 6e-7


def _same_rotations(got: np.ndarray, ref: np.ndarray, tol: float) -> None:
    assert got.shape == ref.shape
    d_pos = np.abs(got - ref).max(1)
    d_neg = np.abs(got + ref).max(1)
    d = np.minimum(d_pos, d_neg)
    # -q is only an alternative where the scalar part is (numerically) zero
    flipped = d_neg < d_pos
    assert np.all(np.abs(ref[flipped, 0]) < 1e-6)
    assert d.max() <= tol, f"max |dq| {d.max():.3e}"


def test_cubochoric_grid(golden):
    _same_rotations(E.so3_cubochoric_grid(6).numpy(), golden["grid_edge6"], 1e-6)


@pytest.mark.parametrize("key,laue,n", [("fz_m3m_2000", 11, 2000), ("fz_6mmm_2000", 9, 2000)])
def test_sample_ori_fz_laue(golden, key, laue, n):
    got = E.sample_ori_fz_laue(laue, n, permute=False).numpy()
    ref = golden[key]
    assert got.shape == ref.shape  # same edge length and the same FZ points kept
    _same_rotations(got, ref, SAMPLER_TOL)
    assert np.all(got[:, 0] >= 0)


def test_sample_ori_fz_laue_angle(golden):
    got = E.sample_ori_fz_laue_angle(11, 8.0, permute=False).numpy()
    _same_rotations(got, golden["fz_angle_m3m_8deg"], SAMPLER_TOL)


def test_sampler_permutation_is_a_permutation():
    a = E.sample_ori_fz_laue(11, 2000, permute=False)
    b = E.sample_ori_fz_laue(11, 2000, permute=True, generator=torch.Generator().manual_seed(5))
    ka = sorted(map(tuple, a.numpy().tolist()))
    kb = sorted(map(tuple, b.numpy().tolist()))
    assert ka == kb and not torch.equal(a, b)


@pytest.mark.parametrize("laue", range(1, 12))
def test_ori_in_fz_laue(golden, laue):
    q = T(golden["fz_in_q"])
    np.testing.assert_array_equal(E.ori_in_fz_laue(q, laue).numpy(), golden[f"fz_in_{laue}"])


@pytest.mark.parametrize("laue", range(1, 12))
def test_laue_elements_as_sets(golden, laue):
    got = E.laue_elements(laue).numpy()
    ref = golden[f"laue_ops_{laue}"]
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got[0], [1.0, 0.0, 0.0, 0.0])
    # the reference's table is float32-rounded literals stored as float64
    d = np.minimum(np.abs(got[:, None] - ref[None]).max(-1), np.abs(got[:, None] + ref[None]).max(-1))
    match = d < 1e-7
    assert np.all(match.sum(1) == 1) and np.all(match.sum(0) == 1), "not the same set of rotations"


def test_disori_angle_laue(golden):
    got = E.disori_angle_laue(T(golden["dis_q1"]), T(golden["dis_q2"]), 11).numpy()
    np.testing.assert_allclose(got, golden["dis_m3m"], rtol=0, atol=2e-6)
    # and in float64 (the product's accurate path) against the float32 reference
    got64 = E.disori_angle_laue(T(golden["dis_q1"]).double(), T(golden["dis_q2"]).double(), 11).numpy()
    np.testing.assert_allclose(got64, golden["dis_m3m"], rtol=0, atol=2e-6)


def test_ori_to_fz_laue_lands_in_fz(golden):
    q = T(golden["fz_in_q"]).double()
    for laue in (9, 11):
        ops = E.laue_elements(laue)
        # the reference's form (laue_fz_ori.py:35-81): q * g with g the
        # operator that gives the largest |w|
        r = E.ori_to_fz_laue(q, laue)
        assert bool(E.ori_in_fz_laue(r, laue).all())
        cand = P.qu_std(P.qu_prod(q[:, None, :], ops[None]))
        brute = cand[torch.arange(q.shape[0]), cand[..., 0].argmax(-1)]
        np.testing.assert_allclose(r.numpy(), brute.numpy(), atol=1e-12)
        # the left form g * q keeps the disorientation class of q
        rl = E.ori_to_fz_laue(q, laue, left=True)
        assert bool(E.ori_in_fz_laue(rl, laue).all())
        # the table holds float32-rounded entries (as the reference's), so
        # the cosine of the half angle is 1 - O(1e-7) and acos maps that to
        # up to ~1e-3 rad
        assert float(E.disori_angle_laue(rl, q, laue).max()) < 2e-3


def test_qu_apply(golden):
    got = E.qu_apply(T(golden["qa_q"]), T(golden["qa_p"])).numpy()
    np.testing.assert_array_equal(got, golden["qa_out"])


@pytest.mark.parametrize("key,shape,tilts,pc,binning", [
    ("geom_60_coords", (60, 60), (0.0, 70.0, 0.0), (0.5, 0.5, 0.6), (1, 1)),
    ("geom_12x16_b2_coords", (12, 16), (3.0, 68.0, 5.0), (0.45, 0.55, 0.7), (2, 2)),
    ("geom_12x16_b1_coords", (12, 16), (3.0, 68.0, 5.0), (0.45, 0.55, 0.7), (1, 1)),
])
def test_geometry_coords(golden, key, shape, tilts, pc, binning):
    g = E.EBSDGeometry(shape, tilts_degrees=tilts, proj_center=pc)
    np.testing.assert_array_equal(g.get_coords_sample_frame(binning).numpy(), golden[key])


def test_local_grid_quats_are_the_cubochoric_map_of_the_grid():
    cu = P.get_local_orientation_grid(2.0, 3)
    q = P.local_orientation_grid_quats(2.0, 3)
    assert q.shape == (343, 4)
    # the centre of the grid is the identity, and angles grow with radius
    np.testing.assert_allclose(q[171].numpy(), [1, 0, 0, 0], atol=1e-7)
    ang = 2 * torch.acos(q[:, 0].clamp(max=1).double())
    rad = cu.double().norm(dim=-1)
    order = torch.argsort(rad)
    assert bool((ang[order][1:] >= 0).all()) and float(ang.max()) < np.deg2rad(2.0) * 3 ** 0.5 * 1.05
