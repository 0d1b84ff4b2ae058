"""Pin the pattern-centre fitter oracle (fitting.py:219-535) and the
product's geometry helpers on the CPU."""

import numpy as np
import torch

from oracle import di_oracle as O
from oracle import fitting_oracle as FO


def test_fitter_geometry_equals_the_golden_pinned_geometry():
    from ebsdtorch_amd.fitting import bruker_geometries_to_SE3, get_3D_pixel_coords_with_periphery, periphery_rays

    for pc, tilt, sec in (((0.5, 0.5, 0.6), 70.0, 0.0), ((0.43, 0.55, 0.71), 65.0, 3.0)):
        R1, t1 = bruker_geometries_to_SE3(torch.tensor([pc]), torch.tensor(-tilt), torch.tensor(sec), (30, 40))
        R2, t2 = O.bruker_geometry_to_SE3(torch.tensor([pc]), torch.tensor(-tilt), torch.tensor(sec), (30, 40))
        assert torch.equal(R1, R2) and torch.equal(t1, t2)
    # without periphery the extended grid is the detector: the same rays as
    # the detector geometry (pinned by the reference's golden vectors)
    pix, pad, shape = get_3D_pixel_coords_with_periphery((30, 40), torch.device("cpu"), 0.5, 0.5, 0.5, 0.5)
    assert pad == (0, 0, 0, 0) and shape == (30, 40)
    R, t = bruker_geometries_to_SE3(torch.tensor([[0.5, 0.5, 0.6]]), torch.tensor(-70.0), torch.tensor(0.0),
                                    (30, 40))
    rays = periphery_rays(pix, R, t)
    torch.testing.assert_close(rays, O.normalized_rays((30, 40), (0.5, 0.5, 0.6)).float(), atol=1e-6, rtol=0)
    pix2, pad2, shape2 = FO.periphery_grid(30, 40, 0.2, 0.9, 0.3, 0.6)
    pix3, pad3, shape3 = get_3D_pixel_coords_with_periphery((30, 40), torch.device("cpu"), 0.2, 0.9, 0.3, 0.6)
    assert pad2 == pad3 and tuple(shape2) == tuple(shape3) and torch.equal(pix2, pix3)


def test_fft_correlation_equals_circular_correlation():
    g = torch.Generator().manual_seed(0)
    h, w = 6, 8
    sim = torch.randn(3, 1, h, w, generator=g, dtype=torch.float64)
    pat = torch.randn(1, 2, h, w, generator=g, dtype=torch.float64)
    got = FO.correlate(sim, torch.fft.rfft2(pat).conj(), (h, w))
    for d in range(3):
        for b in range(2):
            for sy in range(h):
                for sx in range(w):
                    ref = sum(float(sim[d, 0, y, x] * pat[0, b, (y - sy) % h, (x - sx) % w])
                              for y in range(h) for x in range(w))
                    # fftshift moves shift (sy, sx) to ((sy + h//2) % h, (sx + w//2) % w)
                    assert abs(float(got[d, b, (sy + h // 2) % h, (sx + w // 2) % w]) - ref) < 1e-9


def test_fitter_recovers_a_translated_pattern_centre():
    from ebsdtorch_amd.synthetic import band_master_pattern

    mp = band_master_pattern(101, laue_id=11, n_families=10, sigma_deg=2.0, seed=0)
    g = torch.Generator().manual_seed(1)
    quats = O.ori_to_fz_laue(torch.nn.functional.normalize(torch.randn(12, 4, generator=g), dim=1), 11)
    H = W = 24
    truth = (0.5 + 2 / W, 0.5 - 1 / H, 0.6)
    rays = O.normalized_rays((H, W), truth).float()
    pats = O.project_dictionary(mp, rays, quats[:3]).view(3, 1, H, W)
    pc, q = FO.spatial_fft_pc_fitter(mp, quats, pats, 0.3, 0.7, 0.3, 0.7, 0.5, 0.7, 3, batch_size=16)
    assert abs(float(pc[2]) - 0.6) < 1e-6
    # the reference's shift-to-centre arithmetic (fitting.py:524-525) reports
    # the centre mirrored about 0.5 in x and y relative to its own detector
    # geometry (geometry.py:31-92); restated as is
    assert abs(float(pc[0]) - (1 - truth[0])) < 1e-6 and abs(float(pc[1]) - (1 - truth[1])) < 1e-6
    assert torch.equal(q, quats[:3])


def test_cmif_matches_reference_and_brute_force(golden):
    """The oracle's cross mutual information map against the reference's own
    cmif output (bins 4 + 4, per-shift marginals, a source foreground mask)
    and against direct counting of the pixel pairs at every shift."""
    from oracle import fitting_oracle as FO

    dst, src, ind = (torch.tensor(golden[k]) for k in ("cmif_dst", "cmif_src", "cmif_ind"))
    got = FO.cmif(dst, src, ind)
    np.testing.assert_allclose(got.numpy(), golden["cmif_out"], rtol=1e-5, atol=2e-6)
    brute, ones = FO.cmif_brute(dst[:2], src, ind)
    keep = ~ones
    assert int(keep.sum()) > 100
    np.testing.assert_allclose(got[:2].double()[keep].numpy(), brute[keep].numpy(), rtol=1e-4, atol=1e-5)
