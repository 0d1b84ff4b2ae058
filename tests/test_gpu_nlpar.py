"""GPU parity of NLPAR (ebsd_nlpar_distances, ebsd_nlpar_average and
preprocessing.nlpar) against the CPU oracle."""

import pytest
import torch

from oracle import nlpar_oracle as N

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _scan(hs, ws, h, w, seed):
    g = torch.Generator().manual_seed(seed)
    grains = torch.rand(4, h, w, generator=g)
    yy, xx = torch.meshgrid(torch.arange(hs), torch.arange(ws), indexing="ij")
    label = ((xx * 3) // ws + (yy * 2 // hs) * 2).clamp(max=3)
    return grains[label] + 0.1 * torch.randn(hs, ws, h, w, generator=g)


@pytest.mark.parametrize("hs,ws,h,w,k", [(7, 9, 12, 12, 1), (20, 33, 15, 17, 2), (9, 18, 60, 60, 3),
                                         (13, 5, 8, 8, 5)])
def test_nlpar_matches_oracle(hs, ws, h, w, k):
    from ebsdtorch_amd.preprocessing import nlpar, nlpar_distances

    x = _scan(hs, ws, h, w, hs + k)
    print("nlpar", hs, ws, h, w, k, flush=True)
    d = nlpar_distances(x.to(DEV), k).cpu().double()
    d_ref, outside = N.distances(x, k)
    scale = float(d_ref.max())
    torch.testing.assert_close(d, d_ref.double(), atol=2e-6 * scale, rtol=1e-5)
    out = nlpar(x.to(DEV), k_rad=k).cpu()
    ref = N.nlpar(x, k_rad=k)
    torch.testing.assert_close(out, ref, atol=1e-4, rtol=1e-4)
    # the average pass alone, on the oracle's weights: the same sums in order
    from ebsdtorch_amd import _lib
    from ebsdtorch_amd.preprocessing import nlpar_offsets

    w_ref = N.weights(d_ref, outside, h * w, k).contiguous()
    got = torch.empty(hs, ws, h, w, device=DEV)
    offs = nlpar_offsets(k).to(DEV, torch.int32).contiguous()
    xd, wd = x.to(DEV).contiguous(), w_ref.to(DEV)
    _lib.check(_lib.load().ebsd_nlpar_average(xd.data_ptr(), hs, ws, h * w, k, offs.data_ptr(), wd.data_ptr(),
                                              got.data_ptr(), _lib.stream_ptr(DEV)), "ebsd_nlpar_average")
    torch.testing.assert_close(got.cpu(), N.average(x, w_ref, k), atol=2e-6, rtol=2e-6)


def test_do_nlpar_on_experiment_patterns():
    from ebsdtorch_amd.experiment_pats import ExperimentPatterns

    x = _scan(6, 8, 10, 10, 3)
    e = ExperimentPatterns(x.to(DEV))
    e.do_nlpar(k_rad=2)
    torch.testing.assert_close(e.patterns.cpu().view(6, 8, 10, 10), N.nlpar(x, k_rad=2), atol=1e-4, rtol=1e-4)
