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
