"""GPU parity of the PCA path: covariance update, projection GEMM, and
PCA dictionary indexing against the oracle and the reference's golden run."""

import numpy as np
import pytest
import torch

import ebsdtorch_amd as E
from ebsdtorch_amd import kernels, synthetic
from ebsdtorch_amd._lib import DTYPE_BF16, DTYPE_F32
from oracle import di_oracle as O

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("d,batches", [(20, [100, 100, 100]), (200, [513, 1, 77, 300]), (300, [1000])])
def test_cov_update_matches_numpy_cov(d, batches):
    _say(f"cov update d={d} batches={batches}")
    g = torch.Generator().manual_seed(d)
    xs = [torch.randn(b, d, generator=g) * torch.linspace(0.5, 2, d) + torch.linspace(-1, 1, d) for b in batches]
    oc = E.OnlineCovMatrix(d, device=DEV)
    for x in xs:
        oc(x.to(DEV))
    allx = torch.cat(xs).double().numpy()
    ref = np.cov(allx.T)
    got = oc.get_covmat().double().cpu().numpy()
    assert np.abs(got - ref).max() < 2e-5 * np.abs(ref).max()
    assert torch.allclose(oc.mean.double().cpu(), torch.tensor(allx.mean(0)), atol=1e-5)
    assert int(oc.obs.item()) == sum(batches)


def test_cov_update_matches_reference_golden(golden):
    _say("cov update golden")
    x = torch.tensor(golden["cov_x"])
    oc = E.OnlineCovMatrix(20, device=DEV)
    for s in range(0, 300, 100):
        oc(x[s:s + 100].to(DEV))
    assert np.abs(oc.get_covmat().cpu().numpy() - golden["cov_out"]).max() < 1e-5


@pytest.mark.parametrize("dtype,tol", [(DTYPE_F32, 2e-5), (DTYPE_BF16, 2e-5)])
def test_gemm_nt_matches_float64(dtype, tol):
    _say(f"gemm dtype={dtype}")
    g = torch.Generator().manual_seed(9)
    a = torch.randn(300, 1000, generator=g)
    b = torch.randn(130, 1000, generator=g)
    A = kernels.prep_patterns(a.to(DEV), dtype, zero_mean=False)
    B = kernels.prep_patterns(b.to(DEV), dtype, zero_mean=False)
    c = kernels.gemm_nt(A, B).cpu().double()
    ref = A.to_float().cpu().double() @ B.to_float().cpu().double().T
    assert (c - ref).abs().max() < tol * ref.abs().max()
    # accumulate form: C = 2 A B^T - C
    c0 = torch.ones(300, 130, device=DEV)
    kernels.gemm_nt(A, B, out=c0, alpha=2.0, beta=-1.0)
    assert (c0.cpu().double() - (2 * ref - 1)).abs().max() < 2 * tol * ref.abs().max()


def test_pca_driver_matches_reference_golden(golden):
    _say("golden PCA DI")
    mp = E.MasterPattern(torch.tensor(golden["mp33"]).to(DEV), laue_group=11)
    geom = E.EBSDGeometry((8, 8), proj_center=(0.5, 0.5, 0.6))
    exp = E.ExperimentPatterns(torch.tensor(golden["di_exp"]).to(DEV))
    dico = torch.tensor(golden["pca_dict"])
    E.pca_dictionary_index_orientations(mp, geom, exp, n_pca_components=20, experiment_chunk_size=5,
                                        top_k_matches=2, match_dtype=torch.float32,
                                        learn_orientations=[dico], index_orientations=[dico],
                                        component_slice="reference")
    q, m = exp.raw_indexing_results[0]
    ref_q = torch.tensor(golden["pca_quats"])
    ref_m = torch.tensor(golden["pca_metrics"])
    assert torch.allclose(m.cpu(), ref_m, rtol=1e-3, atol=1e-3)
    same = (q.cpu()[:, 0] == ref_q[:, 0]).all(-1)
    assert torch.all(same | ((m.cpu()[:, 0] - m.cpu()[:, 1]).abs() < 1e-4))


def test_pca_indexing_against_fp64_oracle():
    """PCA DI (1600-pixel detector, 300 components) against the oracle's
    PCA DI with an FP64 covariance and eigensolver: top-1 equal except ties
    within 1e-6 in cosine units, measured in the oracle's FP64 PCA space
    (|s_a - s_b| / (|P q| max(|P d_a|, |P d_b|)))."""
    _say("PCA vs FP64 oracle")
    mp = synthetic.band_master_pattern(151, 11, 40, 2.0, seed=0)
    coords = O.detector_coords((40, 40), proj_center=(0.5, 0.5, 0.6))
    learn = O.sample_ori_fz_laue(11, 20000)
    ori = synthetic.random_orientations(500, seed=21)
    rays = O.normalized_rays((40, 40), (0.5, 0.5, 0.6))
    raw = O.interpolate(mp, O.qu_apply(ori[:, None], rays[None]), normalize_coords=False)
    g = torch.Generator().manual_seed(5)
    exp = torch.poisson(200 * raw / raw.mean(1, keepdim=True), generator=g) / 200
    ref_idx, ref_val = O.pca_dictionary_index(mp, coords, exp, learn, learn, 300, top_k=2,
                                              dictionary_chunk_size=4096, experiment_chunk_size=4096,
                                              covmat_dtype=torch.float64)
    eng = E.PCADictionaryIndexer(mp.to(DEV), coords.to(DEV), learn.to(DEV), 300)
    val, idx = eng.index(eng.prepare(exp.to(DEV)), 2)
    got = idx[:, 0].long().cpu()
    diff = torch.nonzero(got != ref_idx[:, 0]).flatten().tolist()
    if diff:
        cov = O.OnlineCovMatrix(coords.shape[0], covmat_dtype=torch.float64, delta_dtype=torch.float64)
        for ob in torch.split(learn, 4096):
            p = O.interpolate(mp, O.qu_apply(ob[:, None, :], coords[None]), normalize_coords=True).view(len(ob), -1)
            cov.update((p - p.mean(-1, keepdim=True)).double())
        basis = O.pca_components(cov.get_eigenvectors(), 300)
        for n in diff:
            cand = learn[torch.tensor([int(got[n]), int(ref_idx[n, 0])])]
            p = O.interpolate(mp, O.qu_apply(cand[:, None, :], coords[None]), normalize_coords=True).view(2, -1)
            pc = (p - p.mean(-1, keepdim=True)).double() @ basis
            qc = (exp[n].double() - exp[n].double().mean()) @ basis
            gap = float(abs(qc @ pc[0] - qc @ pc[1]) / (qc.norm() * pc.norm(dim=-1).max()))
            assert gap < 1e-6, (n, int(got[n]), int(ref_idx[n, 0]), gap)
    _say(f"PCA top-1: {len(diff)} of 500 differ from the FP64 oracle, all ties")
    torch.testing.assert_close(val[:, 0].cpu(), ref_val[:, 0], rtol=1e-5, atol=1e-5)
