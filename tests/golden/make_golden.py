"""Write the reference golden vectors used to pin the CPU oracle.

Run only where the ebsdtorch v0.0.64 sources sit at /root/reference (the
machines that run the GPU tests do not have them):

    python tests/golden/make_golden.py

It imports the side-effect-free reference modules (no network, no demo
script) and records inputs and outputs of the functions on the DI path into
tests/golden/ref_vectors.npz. The DI drivers live in ebsd/indexing.py, whose
module tail is a demo script that imports kikuchipy and downloads data; only
the text of its two driver functions (lines 1-476) is executed here, never the
tail. Inputs are generated with numpy's seeded generator.
"""

from __future__ import annotations

import os
import sys

import numpy as np
import torch

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.npz")


def load_reference():
    sys.path.insert(0, REF)
    from ebsdtorch.s2_and_so3 import sampling, laue_fz_ori, quaternions, sphere, orientations, laue_generators
    from ebsdtorch.ebsd import geometry, master_pattern, experiment_pats
    from ebsdtorch.utils import knn, pca_covmat
    from ebsdtorch.preprocessing import nlpar, clahe
    from ebsdtorch.ebsd import refining
    from ebsdtorch.utils import mutual_information

    src = open(os.path.join(REF, "ebsdtorch/ebsd/indexing.py")).read().splitlines()
    # the drivers end at line 476; the demo script starts at line 479
    ns: dict = {}
    exec(compile("\n".join(src[:476]) + "\n", os.path.join(REF, "ebsdtorch/ebsd/indexing.py"), "exec"), ns)
    return dict(sampling=sampling, laue_fz_ori=laue_fz_ori, quaternions=quaternions, sphere=sphere,
                orientations=orientations, laue_generators=laue_generators, geometry=geometry,
                master_pattern=master_pattern, experiment_pats=experiment_pats, knn=knn,
                pca_covmat=pca_covmat, indexing=ns, nlpar=nlpar, clahe=clahe, refining=refining,
                mutual_information=mutual_information)


def unit_quats(rng, n):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q[q[:, 0] < 0] *= -1
    return q.astype(np.float32)


def unit_vecs(rng, n):
    v = rng.standard_normal((n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v.astype(np.float32)


def band_master_pattern(R, rng, size, n_bands=6):
    """Small smooth master pattern from Gaussian bands (inputs only)."""
    S = size
    X = (2 * np.arange(S) + 1) / S - 1
    yy, xx = np.meshgrid(X, X, indexing="ij")
    sq = torch.tensor(np.stack([xx, yy], -1), dtype=torch.float64)
    nh = R["sphere"].inv_rosca_lambert(sq).numpy()
    # the square's centre (odd sizes) is 0/0 in the inverse map: the pole
    nh[np.isnan(nh).any(-1)] = (0.0, 0.0, 1.0)
    normals = unit_vecs(rng, n_bands).astype(np.float64)
    amps = rng.uniform(0.5, 1.0, n_bands)
    sig = np.deg2rad(8.0)

    def inten(p):
        d = np.abs(p @ normals.T)
        return (amps * np.exp(-np.arcsin(np.clip(d, 0, 1)) ** 2 / (2 * sig ** 2))).sum(-1)

    sh = nh * np.array([1, 1, -1])
    mp = np.concatenate([inten(nh), inten(sh)], axis=1).astype(np.float32)
    mp = (mp - mp.min()) / (1e-4 + mp.max() - mp.min())
    return mp.astype(np.float32)


def main():
    R = load_reference()
    rng = np.random.default_rng(20240917)
    out = {}
    cpu = torch.device("cpu")

    # sampler: cubochoric grid and FZ samples (laue_fz_ori.py:472-513)
    out["grid_edge6"] = R["sampling"].so3_cubochoric_grid(6, device=cpu).numpy()
    out["fz_m3m_2000"] = R["laue_fz_ori"].sample_ori_fz_laue(11, 2000, cpu, permute=False).numpy()
    out["fz_6mmm_2000"] = R["laue_fz_ori"].sample_ori_fz_laue(9, 2000, cpu, permute=False).numpy()
    out["fz_angle_m3m_8deg"] = R["laue_fz_ori"].sample_ori_fz_laue_angle(11, 8.0, cpu, permute=False).numpy()

    # FZ membership for all Laue groups (laue_fz_ori.py:144-239)
    q = unit_quats(rng, 1024)
    out["fz_in_q"] = q
    for lid in range(1, 12):
        out[f"fz_in_{lid}"] = R["laue_fz_ori"].ori_in_fz_laue(torch.tensor(q), lid).numpy()
        out[f"laue_ops_{lid}"] = R["laue_generators"].laue_elements(lid).numpy()

    # disorientation (laue_fz_ori.py:414-469)
    q1, q2 = unit_quats(rng, 64), unit_quats(rng, 64)
    out["dis_q1"], out["dis_q2"] = q1, q2
    out["dis_m3m"] = R["laue_fz_ori"].disori_angle_laue(torch.tensor(q1), torch.tensor(q2), 11, 11).numpy()

    # geometry (geometry.py:186-503)
    g = R["geometry"].EBSDGeometry((60, 60), proj_center=(0.5, 0.5, 0.6))
    out["geom_60_coords"] = g.get_coords_sample_frame((1, 1)).detach().numpy()
    g2 = R["geometry"].EBSDGeometry((12, 16), tilts_degrees=(3.0, 68.0, 5.0), proj_center=(0.45, 0.55, 0.7))
    out["geom_12x16_b2_coords"] = g2.get_coords_sample_frame((2, 2)).detach().numpy()
    out["geom_12x16_b1_coords"] = g2.get_coords_sample_frame((1, 1)).detach().numpy()

    # qu_apply (quaternions.py:186-213)
    qa, pa = unit_quats(rng, 256), unit_vecs(rng, 256)
    out["qa_q"], out["qa_p"] = qa, pa
    out["qa_out"] = R["quaternions"].qu_apply(torch.tensor(qa), torch.tensor(pa)).numpy()

    # side-by-side Lambert (sphere.py:145-185)
    pts = unit_vecs(rng, 4096)
    special = np.array([[0, 0, 1], [0, 0, -1], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0],
                        [0.6, 0.8, 0], [0.6, -0.8, 0.0], [0.5, 0.5, 0.70710678]], dtype=np.float32)
    pts = np.concatenate([pts, special])
    out["lam_pts"] = pts
    out["lam_out"] = R["sphere"].rosca_lambert_side_by_side(torch.tensor(pts)).numpy()

    # MasterPattern.interpolate (master_pattern.py:141-209)
    mp = band_master_pattern(R, rng, 33)
    out["mp33"] = mp
    MP = R["master_pattern"].MasterPattern(torch.tensor(mp), laue_group=11)
    ip = unit_vecs(rng, 3000)
    out["interp_pts"] = ip
    out["interp_out"] = MP.interpolate(torch.tensor(ip), mode="bilinear", align_corners=False,
                                       normalize_coords=False).numpy()

    # ChunkedKNN (knn.py:106-287), FP32, no quantisation
    data = rng.standard_normal((300, 40)).astype(np.float32)
    query = rng.standard_normal((50, 40)).astype(np.float32)
    kn = R["knn"].ChunkedKNN(300, 50, 4, cpu, "angular", torch.float32, False)
    for s in range(0, 300, 64):
        kn.set_data_chunk(torch.tensor(data[s:s + 64]))
        for qs in range(0, 50, 16):
            kn.query_chunk(torch.tensor(query[qs:qs + 16]), qs)
    ki, kd = kn.retrieve_topk()
    out["knn_data"], out["knn_query"] = data, query
    out["knn_idx"], out["knn_val"] = ki.numpy(), kd.numpy()

    # OnlineCovMatrix (pca_covmat.py:94-158)
    cx = (rng.standard_normal((300, 20)) * rng.uniform(0.5, 2, 20) + rng.uniform(-1, 1, 20)).astype(np.float32)
    oc = R["pca_covmat"].OnlineCovMatrix(20)
    for s in range(0, 300, 100):
        oc(torch.tensor(cx[s:s + 100]))
    out["cov_x"] = cx
    out["cov_out"] = oc.get_covmat().numpy()

    # End-to-end DI (indexing.py:14-211), plain FP32
    geom = R["geometry"].EBSDGeometry((8, 8), proj_center=(0.5, 0.5, 0.6))
    pats_q = unit_quats(rng, 16)
    coords = geom.get_coords_sample_frame((1, 1))
    coords = coords / coords.norm(dim=-1, keepdim=True)
    rot = R["quaternions"].qu_apply(torch.tensor(pats_q)[:, None, :], coords[None])
    clean = MP.interpolate(rot, mode="bilinear", align_corners=False, normalize_coords=False).view(16, 8, 8)
    noisy = (clean.numpy() + 0.02 * rng.standard_normal(clean.shape)).astype(np.float32)
    out["di_exp"] = noisy
    for tag, ecs in (("chunked", 5), ("all", 64)):
        torch.manual_seed(1234)
        dico = R["laue_fz_ori"].sample_ori_fz_laue_angle(11, 9.0, cpu, permute=True)
        torch.manual_seed(1234)
        ep = R["experiment_pats"].ExperimentPatterns(torch.tensor(noisy))
        R["indexing"]["dictionary_index_orientations"](
            MP, geom, ep, experiment_chunk_size=ecs, dictionary_resolution_degrees=9.0,
            dictionary_chunk_size=37, top_k_matches=3, match_dtype=torch.float32,
            quantized_if_x86=False)
        out[f"di_{tag}_dict"] = dico.numpy()
        out[f"di_{tag}_quats"] = ep.raw_indexing_results[0][0].numpy()
        out[f"di_{tag}_metrics"] = ep.raw_indexing_results[0][1].numpy()
        out[f"di_{tag}_best"] = ep.orientations.numpy()

    # End-to-end PCA DI (indexing.py:214-476): for 64 features and 20
    # components the reference slice [-20:280] equals the top 20.
    torch.manual_seed(99)
    dico = R["laue_fz_ori"].sample_ori_fz_laue_angle(11, 9.0, cpu, permute=True)
    torch.manual_seed(99)
    ep = R["experiment_pats"].ExperimentPatterns(torch.tensor(noisy))
    R["indexing"]["pca_dictionary_index_orientations"](
        MP, geom, ep, n_pca_components=20, experiment_chunk_size=5,
        dictionary_resolution_learn_deg=9.0, dictionary_resolution_index_deg=9.0,
        dictionary_chunk_size=37, top_k_matches=2, match_dtype=torch.float32, quantized_via_ao=False)
    out["pca_dict"] = dico.numpy()
    out["pca_quats"] = ep.raw_indexing_results[0][0].numpy()
    out["pca_metrics"] = ep.raw_indexing_results[0][1].numpy()

    # NLPAR (preprocessing/nlpar.py:20-192): a 7 x 9 scan of 6 x 5 patterns
    # made of a few smooth clusters plus noise, k_rad 1 and 2
    sy, sx = np.meshgrid(np.arange(7), np.arange(9), indexing="ij")
    base = np.stack([np.sin(0.7 * (i + 1) * np.linspace(0, 3, 30)) for i in range(3)])
    lab = ((sy // 3) + 2 * (sx // 5)) % 3
    scan = base[lab].reshape(7, 9, 6, 5) + 0.3 * rng.standard_normal((7, 9, 6, 5))
    scan = scan.astype(np.float32)
    out["nlpar_in"] = scan
    for k in (1, 2):
        out[f"nlpar_k{k}"] = R["nlpar"].nlpar(torch.tensor(scan), k_rad=k).numpy()
    out["nlpar_k1_bias"] = R["nlpar"].nlpar(torch.tensor(scan), k_rad=1, coeff=0.5,
                                            center_logit_bias=0.3).numpy()

    # CLAHE (preprocessing/clahe.py:18-130): images in [0, 1], shapes that
    # need reflect padding, a clip that redistributes, an unclipped case
    for tag, (b, hh, ww, clip, bins, grid) in {
        "clahe_a": (3, 30, 26, 40.0, 64, (4, 4)),
        "clahe_b": (2, 17, 23, 2.0, 16, (3, 2)),
        "clahe_c": (2, 16, 16, 1000.0, 32, (2, 2)),
    }.items():
        img = rng.random((b, 1, hh, ww)).astype(np.float32) ** 2
        img[0, 0, : hh // 2] *= 0.3  # a dark half: uneven tile histograms
        out[f"{tag}_in"] = img
        out[f"{tag}_out"] = R["clahe"].clahe_grayscale(torch.tensor(img), clip_limit=clip, n_bins=bins,
                                                       grid_shape=grid).numpy()

    # Orientation grid refinement (ebsd/refining.py:71-244): the DI problem
    # above (33x66 MP, 8x8 detector, 16 noisy patterns) started from
    # orientations 1.5 degrees off the truth, FP32 dots; unbinned and binned
    geom8 = R["geometry"].EBSDGeometry((8, 8), proj_center=(0.5, 0.5, 0.6))
    axis = unit_vecs(rng, 16).astype(np.float64)
    ang = np.deg2rad(1.5)
    dq = np.concatenate([np.full((16, 1), np.cos(ang / 2)), axis * np.sin(ang / 2)], 1)
    start = R["quaternions"].qu_prod(torch.tensor(dq, dtype=torch.float32), torch.tensor(pats_q))
    start = start * torch.where(start[:, :1] < 0, -1.0, 1.0)
    out["refine_start"] = start.numpy()
    for tag, binning in (("b1", 1), ("b2", 2)):
        ep = R["experiment_pats"].ExperimentPatterns(torch.tensor(noisy))
        ep.set_orientations(start.clone())
        ep.phases = torch.zeros(16, dtype=torch.uint8)
        MPr = R["master_pattern"].MasterPattern(torch.tensor(mp), laue_group=11)
        dots = R["refining"].orientation_grid_refinement(
            MPr, geom8, ep, grid_semi_edge_in_degrees=2.0, batch_size=4, virtual_binning=binning, n_iter=3,
            kernel_radius_in_steps=1, shrink_factor=0.5, match_dtype=torch.float32)
        out[f"refine_{tag}_ori"] = ep.get_orientations().numpy()
        out[f"refine_{tag}_dots"] = dots[0].numpy()

    # cmif (utils/mutual_information.py:304-673) as the fitter's "mi" metric
    # calls it (fitting.py:486-496): 4 + 4 bins, per-shift marginals;
    # destination (dictionary) images on a full foreground, source
    # (experimental) images zero-padded with their detector area as foreground
    hh, ww = 12, 10
    yy, xx = np.meshgrid(np.linspace(0, 1, hh), np.linspace(0, 1, ww), indexing="ij")
    dst = np.stack([0.5 + 0.45 * np.sin(3 * xx + 2 * yy + ph) for ph in (0.0, 1.0, 2.5)])[:, None]
    dst = np.clip(dst + 0.05 * rng.standard_normal(dst.shape), 0, 0.999).astype(np.float32)
    src = np.zeros((2, 1, hh, ww), np.float32)
    ind = np.zeros((2, 1, hh, ww), np.float32)
    src[:, :, 2:10, 1:8] = np.stack([dst[0, 0, 2:10, 2:9], dst[2, 0, 3:11, 1:8]])[:, None]
    src[:, :, 2:10, 1:8] += 0.05 * rng.standard_normal((2, 1, 8, 7)).astype(np.float32)
    src = np.clip(src, 0, 0.999).astype(np.float32)
    ind[:, :, 2:10, 1:8] = 1.0
    out["cmif_dst"], out["cmif_src"], out["cmif_ind"] = dst, src, ind
    out["cmif_out"] = R["mutual_information"].cmif(images_dst=torch.tensor(dst), images_src=torch.tensor(src),
                                                   src_ind=torch.tensor(ind), bins_dst=4, bins_src=4,
                                                   constant_marginal=False).numpy()

    # MasterPattern.interpolate in every grid_sample mode (master_pattern.py:141-209;
    # refining.py:203-209 and :338-344 call it with align_corners=True, bilinear and bicubic)
    rng2 = np.random.default_rng(77)
    mpm = band_master_pattern(R, rng2, 17)
    mpm = (mpm + 0.2 * rng2.random(mpm.shape)).astype(np.float32)
    out["interp_modes_mp"] = mpm
    pts = unit_vecs(rng2, 1500)
    ang = rng2.uniform(0, 2 * np.pi, 300)
    rim = np.stack([np.cos(ang), np.sin(ang), np.zeros_like(ang)], 1)  # the square's edges
    t = rng2.uniform(-1, 1, 200)
    diag = np.stack([t, t * rng2.choice([-1, 1], 200), rng2.uniform(-1, 1, 200)], 1)
    pts = np.concatenate([pts, rim, diag]).astype(np.float32)
    pts *= rng2.uniform(0.5, 2.0, (len(pts), 1)).astype(np.float32)  # normalize_coords=True rescales
    out["interp_modes_pts"] = pts
    MPm = R["master_pattern"].MasterPattern(torch.tensor(mpm), laue_group=11)
    for mode in ("bilinear", "nearest", "bicubic"):
        for pad in ("zeros", "border", "reflection"):
            for align in (False, True):
                out[f"interp_{mode}_{pad}_{int(align)}"] = MPm.interpolate(
                    torch.tensor(pts), mode=mode, padding_mode=pad, align_corners=align,
                    normalize_coords=True).numpy()

    # Gradient refinement (ebsd/refining.py:247-388): the grid-refinement
    # problem above (33x66 MP, 8x8 detector), 16 patterns from 1.5 degrees
    # off; the last four with inverted contrast (anti-matches: the only
    # patterns whose orientations the function writes back); 10 Adam steps
    geom8 = R["geometry"].EBSDGeometry((8, 8), proj_center=(0.5, 0.5, 0.6))
    gpats = out["di_exp"].copy()
    gpats[12:] = gpats[12:].max() - gpats[12:]
    out["grad_pats"] = gpats
    for tag, md in (("f32", torch.float32), ("f16", torch.float16)):
        ep = R["experiment_pats"].ExperimentPatterns(torch.tensor(gpats))
        ep.set_orientations(torch.tensor(out["refine_start"]).clone())
        ep.phases = torch.zeros(16, dtype=torch.uint8)
        MPg = R["master_pattern"].MasterPattern(torch.tensor(out["mp33"]), laue_group=11)
        R["refining"].orientation_grad_refinement(MPg, geom8, ep, batch_size=8, n_iter=10, match_dtype=md)
        out[f"grad_{tag}_ori"] = ep.get_orientations().numpy()

    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: {len(out)} arrays, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
