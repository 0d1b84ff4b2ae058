"""GPU parity of CLAHE (ebsd_clahe) and per-pattern normalisation
(ebsd_normalize_patterns) against the CPU oracle, and of the
ExperimentPatterns cleaning chain."""

import pytest
import torch

from oracle import preprocess_oracle as PP

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


@pytest.mark.parametrize("n,h,w,grid,nb,clip", [(5, 60, 60, (4, 4), 64, 40.0), (3, 61, 59, (4, 4), 64, 40.0),
                                                (4, 30, 26, (3, 2), 16, 4.0), (2, 64, 64, (8, 8), 32, 0.0),
                                                (3, 120, 120, (4, 4), 64, 40.0)])
def test_clahe_matches_oracle(n, h, w, grid, nb, clip):
    from ebsdtorch_amd.preprocessing import clahe_grayscale

    g = torch.Generator().manual_seed(h + nb)
    x = torch.rand(n, h, w, generator=g) ** 2
    print("clahe", n, h, w, grid, nb, clip, flush=True)
    got = clahe_grayscale(x.to(DEV), clip_limit=clip, n_bins=nb, grid_shape=grid).cpu()
    ref = PP.clahe_grayscale(x, clip_limit=clip, n_bins=nb, grid_shape=grid)
    torch.testing.assert_close(got, ref, atol=2e-5, rtol=0)


@pytest.mark.parametrize("mode", ["minmax", "zeromean", "standard"])
def test_normalize_matches_oracle(mode):
    from ebsdtorch_amd.preprocessing import normalize_patterns

    g = torch.Generator().manual_seed(7)
    x = 3 * torch.rand(9, 3600, generator=g) - 1
    got = normalize_patterns(x.to(DEV).contiguous(), mode).cpu()
    torch.testing.assert_close(got, PP.normalize_per_pattern(x, mode), atol=2e-6, rtol=2e-6)


def test_standard_clean_chain():
    from ebsdtorch_amd.experiment_pats import ExperimentPatterns

    g = torch.Generator().manual_seed(8)
    x = torch.rand(6, 40, 40, generator=g) + torch.linspace(0, 1, 40)[None, None, :]
    e = ExperimentPatterns(x.to(DEV))
    e.standard_clean()
    ref = x - x.mean(0, keepdim=True)
    ref = PP.normalize_per_pattern(ref.view(6, -1), "minmax").view(6, 40, 40)
    ref = PP.clahe_grayscale(ref)
    ref = PP.normalize_per_pattern(ref.view(6, -1), "minmax").view(6, 40, 40)
    torch.testing.assert_close(e.patterns.cpu(), ref, atol=1e-4, rtol=0)
