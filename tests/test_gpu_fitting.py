"""GPU parity of the spatial-FFT pattern-centre fitter against the CPU
oracle (fitting.py:359-535, metric "dot")."""

import pytest
import torch

from oracle import di_oracle as O
from oracle import fitting_oracle as FO

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


@pytest.mark.parametrize("truth,batch", [((0.5 + 2 / 24, 0.5 - 1 / 24, 0.6), 16), ((0.45, 0.55, 0.55), 5)])
def test_spatial_fft_pc_fitter_matches_oracle(truth, batch):
    from ebsdtorch_amd.fitting import spatial_fft_pc_fitter
    from ebsdtorch_amd.master_pattern import MasterPattern
    from ebsdtorch_amd.synthetic import band_master_pattern

    mp = band_master_pattern(101, laue_id=11, n_families=10, sigma_deg=2.0, seed=0)
    g = torch.Generator().manual_seed(1)
    quats = O.ori_to_fz_laue(torch.nn.functional.normalize(torch.randn(12, 4, generator=g), dim=1), 11)
    H = W = 24
    rays = O.normalized_rays((H, W), truth).float()
    pats = O.project_dictionary(mp, rays, quats[[0, 6, 11]]).view(3, 1, H, W)
    print("pc fitter", truth, batch, flush=True)
    pc, q = spatial_fft_pc_fitter(MasterPattern(mp.to(DEV), laue_group=11), pats.clone().to(DEV), 0.3, 0.7, 0.3,
                                  0.7, 0.5, 0.7, 3, dictionary_indexing_batch_size=batch, quats_dict=quats.to(DEV))
    ref_pc, ref_q = FO.spatial_fft_pc_fitter(mp, quats, pats, 0.3, 0.7, 0.3, 0.7, 0.5, 0.7, 3, batch_size=batch)
    torch.testing.assert_close(pc.cpu(), ref_pc, atol=1e-6, rtol=0)
    assert torch.equal(q.cpu(), ref_q)


def test_mi_maps_match_reference_cmif(golden):
    """The "mi" metric's GPU route (indicator and level images, FFT products,
    ebsd_cmif_maps) on the reference's cmif fixture. Shifts where a histogram
    bin holds exactly one pair are excluded: there cmif's "count below 1 ->
    0" rule meets FFT round-off, and rocFFT and the CPU FFT may round the
    same count to either side of 1."""
    from ebsdtorch_amd.fitting import _levels, _mi_maps

    dst, src, ind = (torch.tensor(golden[k]) for k in ("cmif_dst", "cmif_src", "cmif_ind"))
    h, w = src.shape[-2:]
    src_fft = torch.fft.rfft2(torch.cat((ind, _levels(src)), 1).to(DEV)).conj()
    got = _mi_maps(dst.to(DEV), src_fft, (h, w)).cpu()
    ref = torch.tensor(golden["cmif_out"])
    _, ones = FO.cmif_brute(dst, src, ind)
    keep = ~ones
    assert got.shape == ref.shape and int(keep.sum()) > 150
    torch.testing.assert_close(got[keep], ref[keep], rtol=1e-4, atol=1e-5)


def test_mi_maps_odd_width_against_brute_force(golden):
    """An odd image width (12 x 9): the maps equal direct counting at every
    shift (shifts where a histogram bin holds exactly one pair excluded, as
    above)."""
    from ebsdtorch_amd.fitting import _levels, _mi_maps

    dst, src, ind = (torch.tensor(golden[k])[..., :9].contiguous() for k in ("cmif_dst", "cmif_src", "cmif_ind"))
    h, w = src.shape[-2:]
    src_fft = torch.fft.rfft2(torch.cat((ind, _levels(src)), 1).to(DEV)).conj()
    got = _mi_maps(dst.to(DEV), src_fft, (h, w)).cpu()
    ref, ones = FO.cmif_brute(dst, src, ind)
    assert got.shape == ref.shape == (3, 2, h, w)
    keep = ~ones
    assert int(keep.sum()) > 100
    torch.testing.assert_close(got[keep], ref[keep].float(), rtol=1e-4, atol=1e-5)


def test_spatial_fft_pc_fitter_mi_matches_oracle():
    """The fitter with metric "mi" (fitting.py:486-496): the same centre and
    orientations as the oracle's mutual-information fit."""
    from ebsdtorch_amd.fitting import spatial_fft_pc_fitter
    from ebsdtorch_amd.master_pattern import MasterPattern
    from ebsdtorch_amd.synthetic import band_master_pattern

    truth = (0.5 + 2 / 24, 0.5 - 1 / 24, 0.6)
    mp = band_master_pattern(101, laue_id=11, n_families=10, sigma_deg=2.0, seed=0)
    g = torch.Generator().manual_seed(1)
    quats = O.ori_to_fz_laue(torch.nn.functional.normalize(torch.randn(12, 4, generator=g), dim=1), 11)
    H = W = 24
    rays = O.normalized_rays((H, W), truth).float()
    pats = O.interpolate(mp, O.qu_apply(quats[[0, 6, 11]][:, None], rays[None]), normalize_coords=False)
    pats = pats.view(3, 1, H, W)
    print("pc fitter mi", flush=True)
    pc, q = spatial_fft_pc_fitter(MasterPattern(mp.to(DEV), laue_group=11), pats.clone().to(DEV), 0.3, 0.7, 0.3,
                                  0.7, 0.5, 0.7, 3, dictionary_indexing_batch_size=5, quats_dict=quats.to(DEV),
                                  metric="mi")
    ref_pc, ref_q = FO.spatial_fft_pc_fitter(mp, quats, pats, 0.3, 0.7, 0.3, 0.7, 0.5, 0.7, 3, batch_size=5,
                                             metric="mi")
    torch.testing.assert_close(pc.cpu(), ref_pc, atol=1e-6, rtol=0)
    assert torch.equal(q.cpu(), ref_q)
    # the fit finds the centre the patterns were simulated at (the reported
    # x and y are mirrored about 0.5, as in the reference)
    torch.testing.assert_close(pc.cpu(), torch.tensor([1 - truth[0], 1 - truth[1], truth[2]]), atol=1e-6, rtol=0)
