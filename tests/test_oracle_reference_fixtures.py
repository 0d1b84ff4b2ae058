"""Pin the CPU oracles of the steps either side of DI (NLPAR, CLAHE,
orientation grid refinement) to the reference's own outputs: fixtures written
by tests/golden/make_golden.py from the reference modules
(preprocessing/nlpar.py:20-192, preprocessing/clahe.py:18-130,
ebsd/refining.py:71-244) on small inputs; equal bit for bit."""

import numpy as np
import pytest
import torch

from oracle import di_oracle as O
from oracle import nlpar_oracle as N
from oracle import preprocess_oracle as P

T = torch.tensor


@pytest.mark.parametrize("key,k,kw", [("nlpar_k1", 1, {}), ("nlpar_k2", 2, {}),
                                      ("nlpar_k1_bias", 1, {"coeff": 0.5, "center_logit_bias": 0.3})])
def test_nlpar_matches_reference(golden, key, k, kw):
    got = N.nlpar(T(golden["nlpar_in"]), k_rad=k, **kw).numpy()
    np.testing.assert_array_equal(got, golden[key])


@pytest.mark.parametrize("tag,clip,bins,grid", [("clahe_a", 40.0, 64, (4, 4)), ("clahe_b", 2.0, 16, (3, 2)),
                                                ("clahe_c", 1000.0, 32, (2, 2))])
def test_clahe_matches_reference(golden, tag, clip, bins, grid):
    x = T(golden[f"{tag}_in"])[:, 0]
    got = P.clahe_grayscale(x, clip_limit=clip, n_bins=bins, grid_shape=grid).numpy()
    np.testing.assert_array_equal(got, golden[f"{tag}_out"][:, 0])


@pytest.mark.parametrize("tag,binning", [("b1", 1), ("b2", 2)])
def test_grid_refinement_matches_reference(golden, tag, binning):
    mp = T(golden["mp33"])
    rays = O.normalized_rays((8, 8), (0.5, 0.5, 0.6), binning=(binning, binning))
    ori, dots = O.orientation_grid_refinement(mp, 11, rays, T(golden["di_exp"]), T(golden["refine_start"]),
                                              2.0, 4, virtual_binning=binning, n_iter=3, kernel_radius_in_steps=1,
                                              shrink_factor=0.5, match_dtype=torch.float32)
    np.testing.assert_array_equal(dots.numpy(), golden[f"refine_{tag}_dots"])
    np.testing.assert_array_equal(ori.numpy(), golden[f"refine_{tag}_ori"])
