"""End-to-end dictionary-indexing parity on the GPU against the CPU oracle
(the reference's FP32 DI restated) and the reference's golden vectors.

Bar: the GPU's top-1 dictionary index equals the oracle's, except where the
two candidates are tied within 1e-6 in cosine similarity to the query."""

import numpy as np
import pytest
import torch

import ebsdtorch_amd as E
from ebsdtorch_amd import kernels, synthetic
from oracle import di_oracle as O

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")
TIE_COS = 1e-6


def _say(*a):
    print(*a, flush=True)


def _cos(q, d):
    q = q.double() - q.double().mean(-1, keepdim=True)
    return (q * d.double()).sum(-1) / (q.norm(dim=-1) * d.double().norm(dim=-1))


def _assert_top1_parity(got_idx, ref_idx, exp, dict_pats):
    """dict_pats: oracle zero-mean dictionary patterns (N_dict, D)."""
    diff = torch.nonzero(got_idx != ref_idx).flatten()
    for r in diff.tolist():
        cg = _cos(exp[r], dict_pats[got_idx[r]]).item()
        cr = _cos(exp[r], dict_pats[ref_idx[r]]).item()
        assert abs(cg - cr) < TIE_COS, (r, got_idx[r].item(), ref_idx[r].item(), cg, cr)
    return len(diff)


@pytest.mark.parametrize("tag,ecs", [("chunked", 5), ("all", 64)])
def test_driver_matches_reference_golden(golden, tag, ecs):
    _say(f"golden DI {tag}")
    mp = E.MasterPattern(torch.tensor(golden["mp33"]).to(DEV), laue_group=11)
    geom = E.EBSDGeometry((8, 8), proj_center=(0.5, 0.5, 0.6))
    exp = E.ExperimentPatterns(torch.tensor(golden["di_exp"]).to(DEV))
    dico = torch.tensor(golden[f"di_{tag}_dict"])
    E.dictionary_index_orientations(mp, geom, exp, experiment_chunk_size=ecs, dictionary_chunk_size=37,
                                    top_k_matches=3, match_dtype=torch.float32,
                                    dictionary_orientations=[dico])
    q, m = exp.raw_indexing_results[0]
    ref_q = torch.tensor(golden[f"di_{tag}_quats"])
    ref_m = torch.tensor(golden[f"di_{tag}_metrics"])
    idx = exp.raw_indexing_indices[0].cpu()
    assert torch.allclose(m.cpu(), ref_m, rtol=1e-5, atol=1e-5)
    # top-1 quaternion identical unless the metric is tied
    same = (q.cpu()[:, 0] == ref_q[:, 0]).all(-1)
    assert torch.all(same | ((m.cpu()[:, 0] - m.cpu()[:, 1]).abs() < 1e-6))
    assert torch.equal(exp.orientations.cpu(), q.cpu()[:, 0])
    assert exp.phases.dtype == torch.uint8 and torch.all(exp.phases == 0)
    assert idx.dtype == torch.int64


def test_small_config_top1_parity_with_oracle():
    """Config C1 shape: 201x201 m-3m band MP, ~10k FZ dictionary, 60x60
    detector at PC (0.5, 0.5, 0.6), 1,000 noisy patterns, top-1."""
    _say("C1 parity: building inputs")
    mp = synthetic.band_master_pattern(201, 11, 40, 2.0, seed=0)
    dico = O.sample_ori_fz_laue(11, 10000)
    rays = O.normalized_rays((60, 60), (0.5, 0.5, 0.6))
    ori = synthetic.random_orientations(1000, seed=2)
    clean = O.project_dictionary(mp, rays, ori) + 0.0  # zero-mean clean
    g = torch.Generator().manual_seed(3)
    raw = O.interpolate(mp, O.qu_apply(ori[:, None], rays[None]), normalize_coords=False)
    counts = torch.poisson(200 * raw / raw.mean(1, keepdim=True), generator=g)
    exp = counts / 200 + 0.05 * torch.randn(counts.shape, generator=g)
    _say("C1 parity: oracle DI")
    ref_idx, ref_val = O.dictionary_index(mp, rays, exp, dico, top_k=1, dictionary_chunk_size=4096,
                                          experiment_chunk_size=4096)
    _say("C1 parity: GPU DI")
    eng = E.DictionaryIndexer(mp.to(DEV), rays.to(DEV), dico.to(DEV))
    val, idx = eng.index(eng.prepare(exp.to(DEV)), 1)
    dict_pats = O.project_dictionary(mp, rays, dico, chunk=4096)
    n_diff = _assert_top1_parity(idx[:, 0].cpu().long(), ref_idx[:, 0], exp, dict_pats)
    _say(f"C1 parity: {n_diff} tie swaps of 1000")
    assert torch.allclose(val[:, 0].cpu(), ref_val[:, 0], rtol=1e-4, atol=1e-4)
    del clean


def test_driver_with_signal_mask_and_binning():
    _say("mask + binning")
    mp = synthetic.band_master_pattern(101, 11, 20, 3.0, seed=1)
    geom = E.EBSDGeometry((20, 20), proj_center=(0.5, 0.5, 0.6))
    dico = O.sample_ori_fz_laue(11, 3000)
    ori = synthetic.random_orientations(64, seed=5)
    rays_full = O.normalized_rays((20, 20), (0.5, 0.5, 0.6))
    pats = O.interpolate(mp, O.qu_apply(ori[:, None], rays_full[None]), normalize_coords=False).view(64, 20, 20)
    mask = torch.zeros(10, 10, dtype=torch.bool)
    mask[1:9, 2:9] = True
    exp = E.ExperimentPatterns(pats.to(DEV))
    E.dictionary_index_orientations(E.MasterPattern(mp.to(DEV), laue_group=11), geom, exp, signal_mask=mask,
                                    virtual_binning=2, top_k_matches=2, match_dtype=torch.float32,
                                    dictionary_orientations=[dico])
    # oracle: binned MP, binned geometry, binned patterns, masked
    mpb = torch.nn.functional.avg_pool2d(mp[None, None], 2, stride=2)[0, 0]
    rays_b = O.normalized_rays((20, 20), (0.5, 0.5, 0.6), binning=(2, 2))
    pb = torch.nn.functional.avg_pool2d(pats, 2).reshape(64, -1)
    ref_idx, _ = O.dictionary_index(mpb, rays_b, pb, dico, top_k=2, signal_mask=mask,
                                    dictionary_chunk_size=1000, experiment_chunk_size=1000)
    dict_pats = O.project_dictionary(mpb, rays_b, dico, chunk=1000, signal_mask=mask)
    q = pb[:, mask.flatten()]
    _assert_top1_parity(exp.raw_indexing_indices[0][:, 0].cpu(), ref_idx[:, 0], q, dict_pats)


def test_full_size_self_retrieval_333k():
    """Config C2 dictionary size (333k m-3m orientations, 501x501 MP, 60x60):
    noise-free patterns simulated at dictionary orientations must retrieve
    their own entry (size-independent property of the full-size path)."""
    _say("C2 self retrieval: dictionary")
    mp = synthetic.band_master_pattern(501, 11, 40, 2.0, seed=0, device=DEV)
    dico = E.sample_ori_fz_laue(11, 333_000, DEV, permute=True, generator=torch.Generator().manual_seed(1))
    assert 325_000 < dico.shape[0] < 335_000
    rays = E.EBSDGeometry((60, 60), proj_center=(0.5, 0.5, 0.6)).get_rays().to(DEV)
    eng = E.DictionaryIndexer(mp, rays, dico)
    pick = torch.randperm(dico.shape[0], generator=torch.Generator().manual_seed(7))[:4096].to(DEV)
    raw = kernels.project(mp, rays, dico[pick], zero_mean=False).data[:, :3600]
    _say("C2 self retrieval: index")
    val, idx = eng.index(eng.prepare(raw), 16)
    top1 = idx[:, 0].long()
    hit = top1 == pick
    _say(f"C2 self retrieval: {int((~hit).sum())} of 4096 retrieved another entry")
    if not torch.all(hit):
        # only a symmetry-equivalent entry (the FZ boundary is inclusive, so
        # the grid holds a few equivalent pairs) may win: disorientation ~ 0
        bad = torch.nonzero(~hit).flatten()
        dis = E.disori_angle_laue(dico[pick[bad]].double(), dico[top1[bad]].double(), 11)
        assert torch.all(torch.rad2deg(dis) < 0.05), torch.rad2deg(dis)
    assert torch.all(val[:, 1:] <= val[:, :-1])


@pytest.mark.parametrize("components", ["auto", 128])
def test_pca_compressed_search_keeps_fp32_top1(components):
    """FP8 search on the dictionary's leading principal coordinates, then an
    exact FP32 rescore of the top 16 on the full patterns: the top-1 equals
    plain FP32 DI's (up to score ties) on noisy patterns; the basis is
    orthonormal and its second-moment fraction is as reported."""
    mp = synthetic.band_master_pattern(201, 11, 40, 2.0, seed=0, device=DEV)
    dico = E.sample_ori_fz_laue(11, 30_000, DEV, permute=True, generator=torch.Generator().manual_seed(1))
    rays = E.EBSDGeometry((40, 40), proj_center=(0.5, 0.5, 0.6)).get_rays().to(DEV)
    ori = synthetic.random_orientations(2048, seed=5, device=DEV)
    exp = synthetic.simulate_patterns(mp, rays, ori, mean_counts=200.0, gaussian_sigma=0.05, seed=6)
    _say("pca search", components)
    ref = E.DictionaryIndexer(mp, rays, dico, precision="fp32")
    _, ref_idx = ref.index(ref.prepare(exp), 1)
    eng = E.DictionaryIndexer(mp, rays, dico, precision="fp8", rescore_k=16, search_components=components,
                              explained_variance=0.995)
    k = eng.search_components
    assert k % 64 == 0 and 256 <= k <= 1600
    b = eng.basis_t.data[:k, :1600].double()
    torch.testing.assert_close(b @ b.T, torch.eye(k, dtype=torch.float64, device=DEV), atol=1e-5, rtol=0)
    if components == "auto":
        assert eng.explained >= 0.995
    q, qf = eng.prepare_pair(exp)
    val, idx = eng.index(q, 1, queries_f32=qf)
    dict_pats = ref.dictionary.to_float().cpu()
    n_diff = _assert_top1_parity(idx[:, 0].cpu().long(), ref_idx[:, 0].cpu().long(), exp.cpu(), dict_pats)
    _say(f"pca search k={k} ({eng.explained:.4f}): {n_diff} top-1 differences (ties)")
    # the scores are the exact FP32 ones of the reported entries
    torch.testing.assert_close(val[:, 0], (qf.to_float() * ref.dictionary.to_float()[idx[:, 0].long()]).sum(-1),
                               atol=1e-4, rtol=1e-5)


def test_two_phase_driver_with_cpu_resident_patterns():
    """Two phases (an m-3m and a 6/mmm master pattern), patterns kept on the
    host as the reference usually holds them: per phase the top-1 equals the
    oracle's (ties aside), and the combined result takes, per pattern, the
    phase whose top-1 metric is larger (experiment_pats.py:387-451)."""
    _say("two-phase driver")
    mps = [synthetic.band_master_pattern(101, 11, 30, 2.0, seed=3),
           synthetic.band_master_pattern(101, 9, 30, 2.0, seed=4)]
    dicts = [O.sample_ori_fz_laue(11, 3000), O.sample_ori_fz_laue(9, 3000)]
    geom = E.EBSDGeometry((30, 30), proj_center=(0.5, 0.5, 0.6))
    rays = O.normalized_rays((30, 30), (0.5, 0.5, 0.6))
    ori = synthetic.random_orientations(96, seed=8)
    pats = torch.cat([O.interpolate(mps[0], O.qu_apply(ori[:48, None], rays[None]), normalize_coords=False),
                      O.interpolate(mps[1], O.qu_apply(ori[48:, None], rays[None]), normalize_coords=False)])
    pats = pats + 0.01 * torch.randn(pats.shape, generator=torch.Generator().manual_seed(9))
    exp = E.ExperimentPatterns(pats.view(96, 30, 30))  # host-resident
    E.dictionary_index_orientations([E.MasterPattern(m.to(DEV), laue_group=l) for m, l in zip(mps, (11, 9))],
                                    geom, exp, top_k_matches=2, match_dtype=torch.float32,
                                    dictionary_orientations=dicts)
    ref = [O.dictionary_index(m, rays, pats, d, top_k=1, dictionary_chunk_size=1000, experiment_chunk_size=1000)
           for m, d in zip(mps, dicts)]
    for p in range(2):
        dict_pats = O.project_dictionary(mps[p], rays, dicts[p], chunk=1000)
        _assert_top1_parity(exp.raw_indexing_indices[p][:, 0].cpu(), ref[p][0][:, 0], pats, dict_pats)
    m0 = exp.raw_indexing_results[0][1][:, 0].cpu()
    m1 = exp.raw_indexing_results[1][1][:, 0].cpu()
    want = (m1 > m0).to(torch.uint8)
    assert torch.equal(exp.phases.cpu(), want)
    # each pattern was simulated from one phase; it must be recognised
    assert torch.equal(want, (torch.arange(96) >= 48).to(torch.uint8))
    best = torch.where(want.bool(), m1, m0)
    assert torch.equal(exp.best_metrics.cpu(), best)
    q_sel = torch.where(want.bool()[:, None], exp.raw_indexing_results[1][0][:, 0].cpu(),
                        exp.raw_indexing_results[0][0][:, 0].cpu())
    assert torch.equal(exp.orientations.cpu(), q_sel)


def test_di_on_a_master_pattern_read_from_an_emsoft_file():
    """A master pattern read from an EMsoft-layout HDF5 file (energy-weighted
    hemispheres, space group 225 -> m-3m): DI of noisy patterns simulated
    from it, GPU against the oracle, top-1 equal except ties."""
    import os

    from ebsdtorch_amd.io import read_master_pattern

    _say("DI on an EMsoft file's master pattern")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "emsoft_mp_m3m.h5")
    mpo = read_master_pattern(path)
    assert mpo.laue_group == 11
    mp = mpo.master_pattern
    rays = O.normalized_rays((30, 30), (0.5, 0.5, 0.6))
    dico = O.sample_ori_fz_laue(11, 3000)
    ori = synthetic.random_orientations(128, seed=41)
    raw = O.interpolate(mp, O.qu_apply(ori[:, None], rays[None]), normalize_coords=False)
    g = torch.Generator().manual_seed(42)
    exp = (torch.poisson(300 * raw / raw.mean(1, keepdim=True), generator=g) / 300).contiguous()
    ref_idx, _ = O.dictionary_index(mp, rays, exp, dico, top_k=1, dictionary_chunk_size=1000,
                                    experiment_chunk_size=1000)
    ep = E.ExperimentPatterns(exp.view(128, 30, 30))
    E.dictionary_index_orientations(E.MasterPattern(mp.to(DEV), laue_group=mpo.laue_group),
                                    E.EBSDGeometry((30, 30), proj_center=(0.5, 0.5, 0.6)), ep,
                                    match_dtype=torch.float32, dictionary_orientations=[dico])
    dict_pats = O.project_dictionary(mp, rays, dico, chunk=1000)
    n_diff = _assert_top1_parity(ep.raw_indexing_indices[0][:, 0].cpu(), ref_idx[:, 0], exp, dict_pats)
    _say(f"EMsoft-file DI: {n_diff} tie swaps of 128")
    # most patterns find an orientation near their own (the 3,000-entry grid is coarse)
    dis = torch.rad2deg(E.disori_angle_laue(ep.orientations.cpu().double(), ori.double(), 11))
    assert float((dis < 8.0).float().mean()) > 0.9
