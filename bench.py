"""Benchmark of the dictionary-indexing hot path on one MI355X.

Workload (the metric configuration): a 1M-orientation m-3m fundamental-zone
dictionary (cubochoric grid, 995,328 entries) projected from a 501x501
Laue-symmetric band master pattern onto a 60x60 detector (PC 0.5, 0.5, 0.6),
resident in HBM; batches of 65,536 noisy experimental patterns (Poisson at
mean 200 counts + Gaussian sigma 0.05, seeded uniform random orientations)
are indexed against it, top-16 per pattern.

One step = one batch indexed end to end on the GPU: pattern preparation
(zero mean, cast, pad) + similarity GEMM with fused top-k (+ FP32 rescoring
for the precision forms). Inputs are resident in HBM before timing starts.

Prints one JSON line (rank 0). Usage:
  python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp32|bf16|fp16|fp8]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS = {  # MI355X published dense peaks
    "hbm_GBps": 8000.0,
    "fp32": 157.3, "bf16": 2500.0, "fp16": 2500.0, "fp8": 5000.0,  # TFLOP/s
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--precision", default="fp8", choices=["fp32", "bf16", "fp16", "fp8"])
    p.add_argument("--rescore-k", type=int, default=16)
    p.add_argument("--search-components", default="auto",
                   help="low-precision forms: search the dictionary's coordinates in its leading principal "
                        "directions ('auto': enough for --explained-variance; an int; 0: full patterns), "
                        "then rescore the candidates exactly in FP32")
    p.add_argument("--explained-variance", type=float, default=0.97,
                   help="'auto' components: fraction of the dictionary's second moment to keep (the library "
                        "default is 0.995; 0.97 gives 256 components here, with the FP32 top-1 kept on every "
                        "pattern of the agreement sample)")
    p.add_argument("--dict-size", type=int, default=1_000_000)
    p.add_argument("--batch", type=int, default=65536)
    p.add_argument("--topk", type=int, default=16)
    p.add_argument("--mp-size", type=int, default=501)
    p.add_argument("--detector", type=int, default=60)
    p.add_argument("--agree-sample", type=int, default=16384,
                   help="patterns re-indexed in plain FP32 to measure top-1 agreement")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline work (0 = skip)")
    p.add_argument("--pc-fit", type=int, default=1, help="measure the pattern-centre fitter (0 = skip)")
    p.add_argument("--nlpar", type=int, default=1, help="measure NLPAR denoising of the batch as a scan (0 = skip)")
    p.add_argument("--refine", type=int, default=1, help="measure grid refinement of the indexed batch (0 = skip)")
    p.add_argument("--learned-train", type=int, default=2_000_000,
                   help="training patterns of the learned-proposal form (0 = skip that form)")
    p.add_argument("--learned-batch", type=int, default=65536, help="patterns indexed by the learned form")
    p.add_argument("--workload", default="metric", choices=["metric", "c5"],
                   help="metric: the headline configuration; c5: 6/mmm, 120x120, 1M dictionary, streamed scan")
    p.add_argument("--c5-scan", type=int, default=1 << 20, help="c5: patterns in the streamed scan")
    p.add_argument("--c5-train", type=int, default=2_000_000, help="c5: training patterns of the 6/mmm network")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="first run a 1 GiB device copy (reference kernel for PMC byte counters)")
    return p.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported in this build "
              "(the dictionary-indexing path shards by query batch with no exchange, see the design notes)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    import ebsdtorch_amd as E
    from ebsdtorch_amd import kernels, synthetic
    from ebsdtorch_amd._lib import DTYPE_F32

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if a.workload == "c5":
        print(json.dumps(c5_workload(a, dev)), flush=True)
        return
    if a.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    det = a.detector
    D = det * det

    t0 = time.perf_counter()
    mp = synthetic.band_master_pattern(a.mp_size, 11, 40, 2.0, seed=0, device=dev)
    dico = E.sample_ori_fz_laue(11, a.dict_size, dev, permute=True, generator=torch.Generator().manual_seed(1))
    geom = E.EBSDGeometry((det, det), proj_center=(0.5, 0.5, 0.6))
    rays = geom.get_rays().to(dev)
    ori = synthetic.random_orientations(a.batch, seed=2, device=dev)
    exp = synthetic.simulate_patterns(mp, rays, ori, mean_counts=200.0, gaussian_sigma=0.05, seed=3)
    torch.cuda.synchronize()
    log(f"inputs ready in {time.perf_counter() - t0:.1f}s: dictionary {dico.shape[0]}, batch {a.batch}, D={D}")

    # ---- dictionary projection (timed separately: once per dictionary) ----
    rk = a.rescore_k if a.precision != "fp32" else 0
    sc = 0 if (a.precision == "fp32" or rk == 0 or a.search_components == "0") else (
        "auto" if a.search_components == "auto" else int(a.search_components))
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    eng = None
    proj_ms = []
    for rep in range(2):
        eng = None
        torch.cuda.empty_cache()
        ev[0].record()
        eng = E.DictionaryIndexer(mp, rays, dico, precision=a.precision, rescore_k=rk, search_components=sc,
                                  explained_variance=a.explained_variance)
        ev[1].record()
        torch.cuda.synchronize()
        proj_ms.append(ev[0].elapsed_time(ev[1]))
    n_dict = eng.size
    search_dim = eng.search_components if eng.search_components else D
    search_cfg = ({"pca_components": eng.search_components, "second_moment_fraction": round(eng.explained, 5),
                   "rescore": f"top-{max(rk, a.topk)} exactly in FP32 on the full {D} pixels"}
                  if eng.search_components else f"full {D} pixels")
    pca_search = bool(eng.search_components)
    # time the projection kernel alone on the dtype it writes (with a PCA
    # search: the FP32 dictionary the basis and the rescoring come from)
    proj_dtype = DTYPE_F32 if eng.search_components else eng.dtype
    proj_out = eng.dictionary_f32 if eng.search_components else eng.dictionary
    pe = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    pe[0].record()
    kernels.project(mp, rays, dico, proj_dtype, out=proj_out, quads=eng.quads)
    pe[1].record()
    torch.cuda.synchronize()
    proj_kernel_ms = pe[0].elapsed_time(pe[1])
    log(f"dictionary build: {proj_ms[-1]:.1f} ms (engine), projection kernel {proj_kernel_ms:.2f} ms"
        + (f"; search on {eng.search_components} principal components ({eng.explained:.4f} of the second moment)"
           if eng.search_components else ""))

    need_f32 = eng.dictionary_f32 is not None

    # kernel-level timing of the similarity call (events on torch's current
    # stream, where the library launches)
    sim_ms = []

    def timed_step():
        q, qf = eng.prepare_pair(exp)
        k_search = max(a.topk, rk) if need_f32 else a.topk
        val = torch.empty((q.rows, k_search), dtype=torch.float32, device=dev)
        idx = torch.empty((q.rows, k_search), dtype=torch.int32, device=dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        kernels.sim_topk(q, eng.dictionary, k_search, val, idx, merge=False)
        e1.record()
        if need_f32:
            val, idx = kernels.rescore(qf, eng.dictionary_f32, idx, a.topk)
        sim_ms.append((e0, e1))
        return val, idx

    for _ in range(a.warmup):
        timed_step()
    torch.cuda.synchronize()
    sim_ms.clear()
    t_start = time.perf_counter()
    for _ in range(a.steps):
        val, idx = timed_step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_per_step = wall * 1000 / a.steps
    sim_avg_ms = sum(e0.elapsed_time(e1) for e0, e1 in sim_ms) / len(sim_ms)
    value = a.batch * a.steps / wall
    log(f"{a.precision}: {ms_per_step:.1f} ms/step, similarity {sim_avg_ms:.1f} ms, {value:.0f} patterns/s")

    # ---- agreement with plain FP32 DI on a sample (also times FP32 DI) ----
    agree = None
    fp32_rate = None
    if a.precision != "fp32" and a.agree_sample > 0:
        ref = E.DictionaryIndexer(mp, rays, dico, precision="fp32")
        ns = min(a.agree_sample, a.batch)
        qs = ref.prepare(exp[:ns])
        ref.index(qs, 1)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rv, ri = ref.index(ref.prepare(exp[:ns]), a.topk)
        torch.cuda.synchronize()
        fp32_rate = ns / (time.perf_counter() - t1)
        agree = float((ri[:, 0] == idx[:ns, 0]).float().mean().item())
        del ref
        log(f"top-1 agreement with FP32 over {ns} patterns: {agree:.5f}; FP32 DI {fp32_rate:.0f} patterns/s")

    flops = 2.0 * a.batch * n_dict * search_dim
    achieved = flops / (sim_avg_ms / 1000) / 1e12
    peak = PEAKS[a.precision]
    out_bytes = 4 if pca_search else {"fp32": 4, "bf16": 2, "fp16": 2, "fp8": 1}[a.precision]
    proj_bytes = n_dict * (D * out_bytes + 16) + 2 * a.mp_size * a.mp_size * 4 + 12 * D
    proj_gbps = proj_bytes / (proj_kernel_ms / 1000) / 1e9

    refinement = None
    if a.refine:
        refinement = refinement_form(a, mp, geom, exp, ori, eng.orientations[idx[:, 0].long()], det, dev)

    nlpar_res = None
    if a.nlpar:
        nlpar_res = nlpar_form(a, exp, det, dev)

    pcfit = None
    if a.pc_fit:
        pcfit = pc_fit_form(a, mp, exp, det, dev)

    learned = None
    if a.learned_train > 0:
        del eng
        torch.cuda.empty_cache()
        learned = learned_form(a, mp, dico, dev)

    # HBM-side bytes per launch of this kernel at this workload, measured
    # with two PMC passes (FETCH_SIZE, WRITE_SIZE) and committed under
    # profiles/ (null when the workload differs from the measured one)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "sim_topk_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            recs = json.load(f)
        for t in (recs if isinstance(recs, list) else [recs]):
            if (t.get("precision") == a.precision and t.get("dictionary") == n_dict and t.get("batch") == a.batch
                    and t.get("detector") == [det, det] and t.get("topk") == a.topk
                    and t.get("search_dim", D) == search_dim):
                traffic = t["traffic_bytes"]

    cpu = None
    if a.cpu_seconds > 0:
        cpu = cpu_baseline(mp.cpu(), rays.cpu(), exp[:1024].cpu(), dico.cpu(), n_dict, a.cpu_seconds)

    res = {
        "metric": "patterns indexed/s vs plain FP32 DI at >=99.9% top-1 match (1M dict, 60x60)",
        "value": round(value, 1),
        "unit": "patterns/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "bf16": "bf16", "fp16": "f16", "fp8": "fp8_e4m3"}[a.precision],
        "data": "synthetic: 501x501 m-3m band master pattern (40 families, 2 deg), noisy patterns "
                "(Poisson 200 + N(0,0.05^2)) at seeded uniform orientations",
        "config": {
            "workload": f"DI {n_dict} m-3m dictionary x {a.batch}-pattern batch, {det}x{det} detector, top-{a.topk}",
            "dictionary": n_dict, "batch": a.batch, "detector": [det, det], "topk": a.topk,
            "precision": a.precision, "rescore_k": rk,
            "search": search_cfg,
        },
        "top1_agreement_vs_fp32": 1.0 if a.precision == "fp32" else agree,
        "agreement_sample": None if a.precision == "fp32" else min(a.agree_sample, a.batch),
        "fp32_di_patterns_per_s": round(value, 1) if a.precision == "fp32" else (
            None if fp32_rate is None else round(fp32_rate, 1)),
        "speedup_vs_fp32_di": None if fp32_rate is None else round(value / fp32_rate, 2),
        "roofline": {
            "kernel": "sim_topk (similarity GEMM + fused top-k)" + (
                " sim_topk_bt_kernel 192x256" if a.precision != "fp32" else " sim_topk_kernel 128x128"),
            "bound": "mfma",
            "achieved": round(achieved, 2),
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": traffic,
            "avg_launch_ms": round(sim_avg_ms, 3),
            "flops_per_launch": flops,
        },
        "projection": {
            "kernel": "ebsd_project_quads", "bound": "hbm", "ms": round(proj_kernel_ms, 3),
            "achieved": round(proj_gbps, 1), "peak": PEAKS["hbm_GBps"], "unit": "GB/s",
            "frac": round(proj_gbps / PEAKS["hbm_GBps"], 4), "bytes": proj_bytes,
        },
        "refinement": refinement,
        "nlpar": nlpar_res,
        "pc_fit": pcfit,
        "learned_proposal": learned,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


def pc_fit_form(a, mp, exp, det, dev):
    """Spatial-FFT pattern-centre fit (fitting.py:359-535, metric "dot") of 8
    patterns of the batch (simulated at centre (0.5, 0.5, 0.6)): pc_z from
    0.5 to 0.7 in 8 steps, centre search +-0.1 (a 72 x 72 extended detector),
    a 3-degree m-3m dictionary; wall time of the call, and the oracle on the
    same inputs on the host cores."""
    import torch

    from ebsdtorch_amd.fitting import spatial_fft_pc_fitter
    from ebsdtorch_amd.master_pattern import MasterPattern
    from ebsdtorch_amd.orientations import sample_ori_fz_laue_angle

    pats = exp[:8].view(8, 1, det, det).clone()
    quats = sample_ori_fz_laue_angle(11, 3.0, dev, permute=False)
    kw = dict(pc_x_min=0.4, pc_x_max=0.6, pc_y_min=0.4, pc_y_max=0.6, pc_z_min=0.5, pc_z_max=0.7, pc_z_steps=8)
    mpo = MasterPattern(mp, laue_group=11)
    spatial_fft_pc_fitter(mpo, pats.clone(), quats_dict=quats, dictionary_indexing_batch_size=1024, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pc, _ = spatial_fft_pc_fitter(mpo, pats.clone(), quats_dict=quats, dictionary_indexing_batch_size=1024, **kw)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cpu = None
    if a.cpu_seconds > 0:
        from oracle import fitting_oracle as FO

        t1 = time.perf_counter()
        ref_pc, _ = FO.spatial_fft_pc_fitter(mp.cpu(), quats.cpu(), pats.cpu(), 0.4, 0.6, 0.4, 0.6, 0.5, 0.7, 8,
                                             batch_size=1024)
        tc = time.perf_counter() - t1
        cpu = {"seconds": round(tc, 2), "cores": torch.get_num_threads(), "kind": "port",
               "same_centre": bool(torch.allclose(ref_pc, pc.cpu(), atol=1e-6))}
    # the reference's defaults: 1-degree dictionary, 16 pc_z steps from 0.3 to 0.7
    q1 = sample_ori_fz_laue_angle(11, 1.0, dev, permute=False)
    kw1 = dict(kw, pc_z_min=0.3, pc_z_max=0.7, pc_z_steps=16)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    pc1, _ = spatial_fft_pc_fitter(mpo, pats.clone(), quats_dict=q1, dictionary_indexing_batch_size=1024, **kw1)
    torch.cuda.synchronize()
    wall1 = time.perf_counter() - t2
    res = {"workload": f"pattern-centre fit of 8 {det}x{det} patterns, {quats.shape[0]} orientations (3 deg), "
                       f"8 pc_z steps, centre range +-0.1",
           "seconds": round(wall, 3), "centre": [round(v, 4) for v in pc.tolist()], "cpu_oracle": cpu,
           "reference_defaults": {"workload": f"{q1.shape[0]} orientations (1 deg), 16 pc_z steps 0.3-0.7",
                                  "seconds": round(wall1, 3), "centre": [round(v, 4) for v in pc1.tolist()],
                                  "dictionary_patterns_per_s": round(q1.shape[0] * 16 / wall1, 1)}}
    log(f"pc fit: {wall:.2f} s, centre {res['centre']}" + (f", CPU oracle {cpu['seconds']} s" if cpu else "")
        + f"; reference defaults ({q1.shape[0]} orientations x 16 pc_z) {wall1:.2f} s, centre {res['reference_defaults']['centre']}")
    return res


def nlpar_form(a, exp, det, dev):
    """NLPAR denoising (preprocessing/nlpar.py, k_rad 3 as
    ExperimentPatterns.do_nlpar) of the batch viewed as a 256 x 256 scan of
    60 x 60 patterns: the whole call, the two pattern passes with their
    algorithmic HBM bytes (each reads the scan once and writes the
    distances / the denoised scan), and the oracle on a 128 x 128 CPU sample."""
    import torch

    from ebsdtorch_amd.preprocessing import nlpar, nlpar_distances

    n = exp.shape[0]
    side = int(round(math.sqrt(n)))
    if side * side != n:
        return None
    scan = exp.view(side, side, det, det)
    k = 3
    nlpar(scan[:32, :32].contiguous(), k_rad=k)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = nlpar(scan, k_rad=k)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    e = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    e[0].record()
    nlpar_distances(scan, k)
    e[1].record()
    torch.cuda.synchronize()
    dist_ms = e[0].elapsed_time(e[1])
    n_off = (2 * k + 1) ** 2
    data_bytes = n * det * det * 4
    dist_bytes = data_bytes + n * n_off * 4
    cpu = None
    if a.cpu_seconds > 0:
        from oracle import nlpar_oracle as N

        sample = scan[:128, :128].cpu()
        t1 = time.perf_counter()
        N.nlpar(sample, k_rad=k)
        tc = time.perf_counter() - t1
        cpu = {"patterns_per_s": round(128 * 128 / tc, 1), "cores": torch.get_num_threads(), "kind": "port",
               "sample": f"128 x 128 scan ({tc:.2f} s)"}
    del out
    # CLAHE (4 x 4 tiles, 64 bins, clip 40) and minmax of every pattern
    from ebsdtorch_amd.preprocessing import clahe_grayscale, normalize_patterns

    flat = exp.view(n, det * det).clone()
    normalize_patterns(flat, "minmax")
    clahe_grayscale(flat[:1024].view(1024, det, det))
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    e[0].record()
    normalize_patterns(flat, "minmax")
    e[1].record()
    clahe_grayscale(flat.view(n, det, det))
    e[2].record()
    torch.cuda.synchronize()
    norm_ms, clahe_ms = e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2])
    del flat
    res = {
        "workload": f"NLPAR k_rad {k} ({n_off} neighbours) of a {side} x {side} scan of {det}x{det} patterns",
        "patterns_per_s": round(n / wall, 1),
        "ms": round(wall * 1000, 2),
        "distance_pass": {"ms": round(dist_ms, 3), "algorithmic_bytes": dist_bytes,
                          "achieved_GBps": round(dist_bytes / (dist_ms / 1000) / 1e9, 1),
                          "hbm_peak_GBps": PEAKS["hbm_GBps"]},
        "cpu_oracle": cpu,
        "clahe": {"ms": round(clahe_ms, 3), "patterns_per_s": round(n / (clahe_ms / 1000), 1),
                  "hbm_GBps": round(2 * data_bytes / (clahe_ms / 1000) / 1e9, 1)},
        "normalize_minmax": {"ms": round(norm_ms, 3), "hbm_GBps": round(2 * data_bytes / (norm_ms / 1000) / 1e9, 1)},
    }
    log(f"nlpar: {n / wall:.0f} patterns/s ({wall * 1000:.1f} ms), distance pass {dist_ms:.2f} ms "
        f"({res['distance_pass']['achieved_GBps']} GB/s); CLAHE {clahe_ms:.2f} ms, minmax {norm_ms:.2f} ms")
    return res


def refinement_form(a, mp, geom, exp, truth, found, det, dev):
    """Orientation grid refinement of the indexed batch (refining.py:71-244
    with its defaults: kernel radius 1 = 27 candidates, 3 iterations, shrink
    0.5; semi-edge 1 degree, no binning, match in FP32): patterns/s of the
    whole call with the FP8-indexed orientations as the start, disorientation
    to the truth before and after, the refinement kernel alone, and the
    oracle on a CPU sample."""
    import torch

    from ebsdtorch_amd.experiment_pats import ExperimentPatterns
    from ebsdtorch_amd.master_pattern import MasterPattern
    from ebsdtorch_amd.orientations import disori_angle_laue
    from ebsdtorch_amd.refining import orientation_grid_refinement

    n = exp.shape[0]
    pats = exp.view(n, det, det)
    mpo = MasterPattern(mp, laue_group=11)
    kw = dict(grid_semi_edge_in_degrees=1.0, batch_size=n, virtual_binning=1, n_iter=3, kernel_radius_in_steps=1)

    def run(q0):
        e = ExperimentPatterns(pats[:q0.shape[0]])
        e.set_orientations(q0.clone())
        e.phases = torch.zeros(q0.shape[0], dtype=torch.uint8, device=dev)
        return e, orientation_grid_refinement(mpo, geom, e, **kw)[0]

    run(found[:4096].contiguous())  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e, dots = run(found)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    before = torch.rad2deg(disori_angle_laue(found.double(), truth.double(), 11))
    after = torch.rad2deg(disori_angle_laue(e.get_orientations().double(), truth.double(), 11))
    qs = torch.tensor([0.5, 0.9, 0.99], dtype=torch.float64, device=dev)
    # one kernel launch over the batch (first iteration)
    from ebsdtorch_amd import kernels
    from ebsdtorch_amd.refining import _grid_levels
    from ebsdtorch_amd.orientations import laue_elements

    rays = geom.get_coords_sample_frame(binning=(1, 1)).view(-1, 3)
    rays = (rays / rays.norm(dim=-1, keepdim=True)).to(dev).contiguous()
    grids = _grid_levels(1.0, 1, 3, 0.5, 3).to(dev)
    x = kernels.prep_patterns(exp, 0, zero_mean=True, unit_norm=False)
    quads = kernels.mp_quad_table(mp, torch.float32)
    ops = laue_elements(11).to(dev, torch.float32)
    lev = torch.zeros(n, dtype=torch.int32, device=dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    kernels.grid_refine_step(quads, rays, x, found, grids, lev, ops)
    e1.record()
    torch.cuda.synchronize()
    step_ms = e0.elapsed_time(e1)
    samples = n * grids.shape[1] * rays.shape[0]
    cpu = None
    if a.cpu_seconds > 0:
        from oracle import di_oracle as O

        ns = 256
        t1 = time.perf_counter()
        O.orientation_grid_refinement(mp.cpu(), 11, rays.cpu(), pats[:ns].cpu(), found[:ns].cpu(), 1.0, ns,
                                      n_iter=3, kernel_radius_in_steps=1)
        tc = time.perf_counter() - t1
        cpu = {"patterns_per_s": round(ns / tc, 2), "cores": torch.get_num_threads(), "kind": "port",
               "sample": f"{ns} patterns ({tc:.1f} s)"}
    out = {
        "workload": f"grid refinement of {n} FP8-indexed patterns {det}x{det}: 27 candidates (radius 1, semi-edge "
                    f"1 deg), 3 iterations, shrink 0.5",
        "patterns_per_s": round(n / wall, 1),
        "ms": round(wall * 1000, 2),
        "disorientation_deg_quantiles_before": [round(v, 3) for v in torch.quantile(before, qs).tolist()],
        "disorientation_deg_quantiles_after": [round(v, 3) for v in torch.quantile(after, qs).tolist()],
        "mean_best_dot_by_iteration": [round(v, 4) for v in dots.mean(0).tolist()],
        "step_kernel": {"ms": round(step_ms, 3), "samples": samples,
                        "gsamples_per_s": round(samples / (step_ms / 1000) / 1e9, 1)},
        "cpu_oracle": cpu,
    }
    log(f"refinement: {n / wall:.0f} patterns/s, median error {out['disorientation_deg_quantiles_before'][0]} -> "
        f"{out['disorientation_deg_quantiles_after'][0]} deg, step kernel {step_ms:.2f} ms")
    return out


def learned_form(a, mp, dico, dev):
    """The learned-proposal configuration: 64x64 detector, network trained
    on `learned_train` patterns simulated on the fly (uniform orientations,
    Poisson mean 100-400 + N(0, 0.05^2), pattern centre jittered +-0.01),
    then `learned_batch` noisy patterns at seeded orientations indexed by
    proposal -> 2,197-orientation local dictionary within 3 degrees, searched
    on 4x4-binned patterns (16x16), its best 8 rescored on the full 64x64
    patterns -> FP8 full-dictionary search (FP32 top-16 rescore) below the
    calibrated threshold. Compared with plain FP32 DI of the same 1M dictionary on a
    sample, and with the ground truth (disorientation)."""
    import torch

    import ebsdtorch_amd as E
    from ebsdtorch_amd import kernels
    from ebsdtorch_amd.learned import (LearnedIndexer, PatternSimulator, ProposalNet, calibrate_threshold,
                                       train_proposal_net)
    from ebsdtorch_amd.orientations import disori_angle_laue

    shape = (64, 64)
    geom = E.EBSDGeometry(shape, proj_center=(0.5, 0.5, 0.6))
    rays = geom.get_rays().to(dev)
    sim = PatternSimulator(mp, geom, mean_counts=(100.0, 400.0), sigma=0.05, pc_jitter=0.01, seed=5)
    net = ProposalNet(shape, dev, laue_id=11, compute="bf16", seed=0)
    tlog = train_proposal_net(net, sim, a.learned_train, batch=512, lr=1e-3, warmup=200, log_every=500)
    log(f"learned: trained on {tlog.patterns} patterns in {tlog.seconds:.1f}s "
        f"({tlog.patterns / tlog.seconds:.0f}/s), loss {tlog.loss_curve[0][1]:.3f} -> {tlog.loss_curve[-1][1]:.4f}")
    sc = 0 if a.search_components == "0" else ("auto" if a.search_components == "auto" else int(a.search_components))
    fb = E.DictionaryIndexer(mp, rays, dico, precision="fp8", rescore_k=16, dict_unit_norm=True, search_components=sc,
                             explained_variance=a.explained_variance)
    binning, rescore_k = 4, 8
    idx = LearnedIndexer(net, mp, rays, 11, fallback=fb, threshold=0.0, binning=binning,
                         local_rays=geom.get_rays((binning, binning)).to(dev), rescore_k=rescore_k)
    cal = PatternSimulator(mp, geom, mean_counts=(100.0, 400.0), sigma=0.05, pc_jitter=0.0, seed=6)
    idx.threshold = calibrate_threshold(idx, cal, n=8192, max_error_deg=1.0, precision=0.999)
    test = PatternSimulator(mp, geom, mean_counts=(100.0, 400.0), sigma=0.05, pc_jitter=0.0, seed=7)
    n = a.learned_batch
    x, q_true = test.batch(n)
    idx.index(x[:4096])  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = idx.index(x)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    rate = n / wall
    # the local search (binned) and the full-resolution rescore, one launch
    # each over the batch
    xb = idx._local_queries(x)
    xf = kernels.PatternMatrix(x.contiguous(), None, 0, x.shape[1])
    q0 = net.propose(x)
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record()
    _, cand = kernels.local_refine(idx.quads, idx.local_rays, xb, q0, idx.grid, topk=rescore_k)
    e1.record()
    kernels.local_rescore(idx.quads, rays, xf, q0, idx.grid, cand)
    e2.record()
    torch.cuda.synchronize()
    local_ms = e0.elapsed_time(e1)
    rescore_ms = e1.elapsed_time(e2)
    samples = n * idx.grid.shape[0] * xb.n_pix
    rescore_samples = n * rescore_k * x.shape[1]
    err = torch.rad2deg(disori_angle_laue(res["orientations"].double(), q_true.double(), 11))
    qs = torch.tensor([0.5, 0.9, 0.99, 0.999], dtype=torch.float64, device=err.device)
    quant = [round(v, 3) for v in torch.quantile(err, qs).tolist()]
    fallback_rate = float(res["fallback"].float().mean())
    # plain FP32 DI of the same dictionary on a sample: rate and agreement
    ns = min(8192, n)
    ref = E.DictionaryIndexer(mp, rays, dico, precision="fp32")
    ref.index(ref.prepare(x[:1024]), 1)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    rv, ri = ref.index(ref.prepare(x[:ns]), 1)
    torch.cuda.synchronize()
    fp32_rate = ns / (time.perf_counter() - t1)
    q_fp32 = ref.orientations[ri[:, 0].long()]
    err_fp32 = torch.rad2deg(disori_angle_laue(q_fp32.double(), q_true[:ns].double(), 11))
    vs_fp32 = torch.rad2deg(disori_angle_laue(res["orientations"][:ns].double(), q_fp32.double(), 11))
    del ref, fb
    torch.cuda.empty_cache()
    out = {
        "workload": f"learned proposal: {n} patterns 64x64, local dictionary {idx.grid.shape[0]} within 3 deg "
                    f"on {binning}x{binning}-binned patterns, best {rescore_k} rescored at 64x64, "
                    f"fallback FP8 DI over {dico.shape[0]}",
        "patterns_per_s": round(rate, 1),
        "fp32_di_patterns_per_s": round(fp32_rate, 1),
        "speedup_vs_fp32_di": round(rate / fp32_rate, 2),
        "fallback_rate": round(fallback_rate, 4),
        "threshold": round(idx.threshold, 5),
        "disorientation_deg_quantiles_vs_truth": dict(zip(["p50", "p90", "p99", "p99.9"], quant)),
        "frac_within_1deg_of_truth": round(float((err <= 1.0).float().mean()), 5),
        "fp32_di_frac_within_1deg_of_truth": round(float((err_fp32 <= 1.0).float().mean()), 5),
        "frac_within_1deg_of_fp32_di": round(float((vs_fp32 <= 1.0).float().mean()), 5),
        "train": {"patterns": tlog.patterns, "seconds": round(tlog.seconds, 2),
                  "patterns_per_s": round(tlog.patterns / tlog.seconds, 1),
                  "loss_first": round(tlog.loss_curve[0][1], 4), "loss_last": round(tlog.loss_curve[-1][1], 4),
                  "compute": "bf16 operands, fp32 master weights"},
        "local_refine_kernel": {"ms": round(local_ms, 3), "samples": samples,
                                "gsamples_per_s": round(samples / (local_ms / 1000) / 1e9, 1)},
        "local_rescore_kernel": {"ms": round(rescore_ms, 3), "samples": rescore_samples,
                                 "gsamples_per_s": round(rescore_samples / (rescore_ms / 1000) / 1e9, 1)},
    }
    log(f"learned: {rate:.0f} patterns/s vs FP32 DI {fp32_rate:.0f}/s, fallback {fallback_rate:.4f}, "
        f"err quantiles {quant}, local kernel {local_ms:.1f} ms, rescore kernel {rescore_ms:.1f} ms")
    return out


def c5_workload(a, dev):
    """Combined stress case: 6/mmm band master pattern (501x501, 60
    families), 120x120 detector, ~1M fundamental-zone dictionary, a scan of
    `c5_scan` noisy patterns (seeded uniform orientations, Poisson mean
    U(100, 400) + N(0, 0.05^2)) simulated and indexed in 65,536-pattern
    chunks by (1) FP8 similarity + FP32 top-16 rescore and (2) the learned
    proposal retrained for 6/mmm with that search as fallback. Cosine
    ranking (unit-norm dictionary) in both, so the forms share one resident
    dictionary; FP32 agreement is measured on a 4,096-pattern sample with the
    FP32 copy of the same dictionary."""
    import torch

    import ebsdtorch_amd as E
    from ebsdtorch_amd import kernels, synthetic
    from ebsdtorch_amd._lib import DTYPE_F32, DTYPE_FP8_E4M3
    from ebsdtorch_amd.learned import (LearnedIndexer, PatternSimulator, ProposalNet, calibrate_threshold,
                                       train_proposal_net)
    from ebsdtorch_amd.orientations import disori_angle_laue

    laue = 9
    det = 120
    D = det * det
    chunk = 65536
    mp = synthetic.band_master_pattern(501, laue, 60, 2.0, seed=10, device=dev)
    dico = E.sample_ori_fz_laue(laue, 1_000_000, dev, permute=True, generator=torch.Generator().manual_seed(11))
    geom = E.EBSDGeometry((det, det), proj_center=(0.5, 0.5, 0.6))
    rays = geom.get_rays().to(dev)
    n_dict = dico.shape[0]
    torch.cuda.synchronize()
    sc = 0 if a.search_components == "0" else ("auto" if a.search_components == "auto" else int(a.search_components))
    tb = time.perf_counter()
    eng = E.DictionaryIndexer(mp, rays, dico, precision="fp8", rescore_k=16, dict_unit_norm=True,
                              search_components=sc, explained_variance=a.explained_variance)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - tb
    search_dim = eng.search_components if eng.search_components else D
    pe = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    pe[0].record()
    if not eng.search_components:
        kernels.project(mp, rays, dico, DTYPE_FP8_E4M3, unit_norm=True, out=eng.dictionary, quads=eng.quads)
    pe[1].record()
    kernels.project(mp, rays, dico, DTYPE_F32, unit_norm=True, out=eng.dictionary_f32, quads=eng.quads)
    pe[2].record()
    torch.cuda.synchronize()
    proj8 = None if eng.search_components else pe[0].elapsed_time(pe[1])
    proj32 = pe[1].elapsed_time(pe[2])
    bytes8 = None if eng.search_components else n_dict * (eng.dictionary.ld + 4 + 16) + 2 * 501 * 501 * 4 + 12 * D
    bytes32 = n_dict * (eng.dictionary_f32.ld * 4 + 16) + 2 * 501 * 501 * 4 + 12 * D
    log(f"c5: dictionary {n_dict} x {D}: built in {build_s:.1f} s, projection FP32 {proj32:.1f} ms"
        + (f", FP8 {proj8:.1f} ms" if proj8 is not None else
           f"; search on {eng.search_components} principal components ({eng.explained:.4f})"))

    sim = PatternSimulator(mp, geom, mean_counts=(100.0, 400.0), sigma=0.05, pc_jitter=0.01, seed=20)
    net = ProposalNet((det, det), dev, laue_id=laue, compute="bf16", seed=0)
    tlog = train_proposal_net(net, sim, a.c5_train, batch=256, lr=1e-3, warmup=200, log_every=1000)
    log(f"c5: trained on {tlog.patterns} patterns in {tlog.seconds:.1f}s, loss {tlog.loss_curve[-1][1]:.4f}")
    # the local search scores 4x4-binned patterns (30x30) against the 2,197
    # candidates and rescores its best 8 at 120x120; the network and the
    # fallback search see 120x120
    c5_binning = 4
    lidx = LearnedIndexer(net, mp, rays, laue, fallback=eng, threshold=0.0, binning=c5_binning,
                          local_rays=geom.get_rays((c5_binning, c5_binning)).to(dev), rescore_k=8)
    lidx.threshold = calibrate_threshold(
        lidx, PatternSimulator(mp, geom, mean_counts=(100.0, 400.0), sigma=0.05, pc_jitter=0.0, seed=21),
        n=8192)
    scan = PatternSimulator(mp, geom, mean_counts=(100.0, 400.0), sigma=0.05, pc_jitter=0.0, seed=22)

    t_di = t_le = 0.0
    sim_ms = []
    err_di, err_le, fb = [], [], []
    agree = None
    n_done = 0
    k = 16
    while n_done < a.c5_scan:
        b = min(chunk, a.c5_scan - n_done)
        x, q_true = scan.batch(b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        q8, q32 = eng.prepare_pair(x, unit_norm=True)
        val = torch.empty((b, k), device=dev)
        idx = torch.empty((b, k), dtype=torch.int32, device=dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        kernels.sim_topk(q8, eng.dictionary, k, val, idx, merge=False)
        e1.record()
        rv, ri = kernels.rescore(q32, eng.dictionary_f32, idx, 1)
        torch.cuda.synchronize()
        t_di += time.perf_counter() - t0
        sim_ms.append((b, e0.elapsed_time(e1)))
        if agree is None:
            ns = 4096
            fv = torch.empty((ns, 1), device=dev)
            fi = torch.empty((ns, 1), dtype=torch.int32, device=dev)
            kernels.sim_topk(q32.narrow(0, ns), eng.dictionary_f32, 1, fv, fi, merge=False)
            agree = float((fi[:, 0] == ri[:ns, 0]).float().mean())
        t0 = time.perf_counter()
        res = lidx.index(x)
        torch.cuda.synchronize()
        t_le += time.perf_counter() - t0
        err_di.append(torch.rad2deg(disori_angle_laue(dico[ri[:, 0].long()].double(), q_true.double(), laue)))
        err_le.append(torch.rad2deg(disori_angle_laue(res["orientations"].double(), q_true.double(), laue)))
        fb.append(res["fallback"])
        n_done += b
        log(f"c5: {n_done}/{a.c5_scan} patterns, FP8 DI {n_done / t_di:.0f}/s, learned {n_done / t_le:.0f}/s")
    err_di = torch.cat(err_di)
    err_le = torch.cat(err_le)
    fb = torch.cat(fb)
    qs = torch.tensor([0.5, 0.9, 0.99, 0.999], dtype=torch.float64, device=err_di.device)
    full = [m for nb, m in sim_ms if nb == sim_ms[0][0]]
    sim_avg = sum(full) / len(full)
    flops = 2.0 * sim_ms[0][0] * n_dict * search_dim
    return {
        "metric": "c5: patterns indexed/s, 6/mmm 1M dictionary, 120x120, streamed scan",
        "workload": f"6/mmm {n_dict} dictionary x {a.c5_scan}-pattern scan in {chunk}-pattern chunks, {det}x{det}",
        "n_gpus": 1,
        "fp8_di": {"patterns_per_s": round(a.c5_scan / t_di, 1), "top1_agreement_vs_fp32": agree,
                   "agreement_sample": 4096,
                   "disorientation_deg_quantiles_vs_truth": dict(zip(["p50", "p90", "p99", "p99.9"],
                                                                     [round(v, 3) for v in torch.quantile(err_di, qs).tolist()])),
                   "frac_within_1deg_of_truth": round(float((err_di <= 1.0).float().mean()), 5)},
        "similarity_kernel": {"avg_launch_ms": round(sim_avg, 2), "achieved_TFLOPs": round(flops / sim_avg / 1e9, 1),
                              "peak_TFLOPs": PEAKS["fp8"], "mfma_frac": round(flops / sim_avg / 1e9 / PEAKS["fp8"], 4)},
        "projection": {"fp8_ms": None if proj8 is None else round(proj8, 2),
                       "fp8_GBps": None if proj8 is None else round(bytes8 / proj8 / 1e6, 1),
                       "fp32_ms": round(proj32, 2), "fp32_GBps": round(bytes32 / proj32 / 1e6, 1),
                       "hbm_peak_GBps": PEAKS["hbm_GBps"]},
        "search": ({"pca_components": eng.search_components, "second_moment_fraction": round(eng.explained, 5),
                    "dictionary_build_s": round(build_s, 2)} if eng.search_components else f"full {D} pixels"),
        "learned_proposal": {"patterns_per_s": round(a.c5_scan / t_le, 1), "fallback_rate": round(float(fb.float().mean()), 4),
                             "local_binning": c5_binning, "rescore_k": lidx.rescore_k,
                             "threshold": round(lidx.threshold, 5),
                             "disorientation_deg_quantiles_vs_truth": dict(zip(["p50", "p90", "p99", "p99.9"],
                                                                               [round(v, 3) for v in torch.quantile(err_le, qs).tolist()])),
                             "frac_within_1deg_of_truth": round(float((err_le <= 1.0).float().mean()), 5),
                             "train": {"patterns": tlog.patterns, "seconds": round(tlog.seconds, 1),
                                       "loss_last": round(tlog.loss_curve[-1][1], 4)}},
    }


def cpu_baseline(mp, rays, exp, dico, n_dict, seconds):
    """The oracle on the host cores, two variants of the reference's CPU DI
    (same chunking as the reference demo: 4096 x 4096), each on 1,024
    patterns against a dictionary slice sized for ~`seconds` of work, scaled
    to the full dictionary: plain FP32 (`value`, the oracle's semantics) and
    the reference's actual CPU default, int8 dynamic quantisation of each
    dictionary chunk (quantized_if_x86, knn.py:52-72; `int8`)."""
    import platform
    import warnings

    import torch

    from oracle import di_oracle as O

    warnings.filterwarnings("ignore", message=".*torch.ao.quantization is deprecated.*")
    cores = torch.get_num_threads()

    def timed(quantized, budget):
        n_sub = 4096
        while True:
            t0 = time.perf_counter()
            O.dictionary_index(mp, rays, exp, dico[:n_sub], top_k=16, dictionary_chunk_size=4096,
                               experiment_chunk_size=4096, quantized=quantized)
            t = time.perf_counter() - t0
            if t >= budget / 3 or n_sub >= n_dict:
                return n_sub, t
            n_sub = min(n_dict, int(n_sub * max(2.0, budget / max(t, 1e-3))))

    n_sub, t = timed(False, seconds)
    rate = exp.shape[0] / (t * n_dict / n_sub)
    log(f"cpu baseline: {exp.shape[0]} patterns x {n_sub} dictionary entries in {t:.1f}s on {cores} threads")
    n8, t8 = timed(True, seconds / 2)
    rate8 = exp.shape[0] / (t8 * n_dict / n8)
    log(f"cpu baseline int8: {exp.shape[0]} patterns x {n8} dictionary entries in {t8:.1f}s")
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    cfg = torch.__config__.show()
    libs = [l.strip(" -") for l in cfg.splitlines() if "Math Kernel Library" in l or "MKL-DNN" in l]
    return {"value": round(rate, 3), "unit": "patterns/s", "cores": cores, "kind": "port",
            "sample": f"{exp.shape[0]} patterns x first {n_sub} of {n_dict} dictionary entries "
                      f"({t:.1f} s), scaled to the full dictionary",
            "int8": {"value": round(rate8, 3), "unit": "patterns/s",
                     "what": "the reference's CPU default: int8 quantize_dynamic of each dictionary chunk",
                     "sample": f"{exp.shape[0]} patterns x first {n8} dictionary entries ({t8:.1f} s), scaled"},
            "host": {"os_cpu_count": os.cpu_count(), "threads_used": cores, "model": model, "libraries": libs}}


if __name__ == "__main__":
    main()
