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
    p.add_argument("--agree", type=int, default=1,
                   help="index the batch (and a held-out set) in plain FP32 too, for agreement (0 = skip, "
                        "for profiling runs)")
    p.add_argument("--agree-sample", type=int, default=None,
                   help="0: the same as --agree 0 --forms 0 (a profiling run of the headline form alone)")
    p.add_argument("--forms", type=int, default=1,
                   help="also measure BF16 and FP8 on full patterns and the library-default search (0 = skip)")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline work (0 = skip)")
    p.add_argument("--pc-fit", type=int, default=1, help="measure the pattern-centre fitter (0 = skip)")
    p.add_argument("--nlpar", type=int, default=1, help="measure NLPAR denoising of the batch as a scan (0 = skip)")
    p.add_argument("--refine", type=int, default=1, help="measure grid refinement of the indexed batch (0 = skip)")
    p.add_argument("--learned-train", type=int, default=2_000_000,
                   help="training patterns of the learned-proposal form (0 = skip that form)")
    p.add_argument("--learned-batch", type=int, default=65536, help="patterns indexed by the learned form")
    p.add_argument("--workload", default="metric", choices=["metric", "c3", "c5", "projection", "library-gemm"],
                   help="metric: the headline configuration; c3: the precision forms and PCA DI at 80x80; "
                        "c5: 6/mmm, 120x120, 1M dictionary, streamed scan")
    p.add_argument("--proj-config", default="metric", choices=["metric", "c3", "c5"],
                   help="projection workload: metric (m-3m, 60x60), c3 (m-3m, 80x80) or c5 (6/mmm, 60 families, "
                        "120x120)")
    p.add_argument("--c3-pca-components", type=int, default=1000, help="c3: components of the PCA DI form")
    p.add_argument("--c3-learn", type=int, default=1 << 20,
                   help="c3: learning-dictionary rows for the PCA DI covariance (first rows of the dictionary)")
    p.add_argument("--c5-scan", type=int, default=1 << 20, help="c5: patterns in the streamed scan")
    p.add_argument("--c5-train", type=int, default=2_000_000, help="c5: training patterns of the 6/mmm network")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="first run a 1 GiB device copy (reference kernel for PMC byte counters)")
    a = p.parse_args()
    if a.agree_sample == 0:
        a.agree, a.forms = 0, 0
    return a


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported "
              "(the dictionary-indexing path would shard by query batch with no exchange between GPUs)",
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
    if a.workload == "c3":
        print(json.dumps(c3_workload(a, dev)), flush=True)
        return
    if a.workload == "projection":
        print(json.dumps(projection_workload(a, dev)), flush=True)
        return
    if a.workload == "library-gemm":
        print(json.dumps(library_gemm_workload(a, dev)), flush=True)
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

    # ---- the headline engine: built once per dictionary (timed) ----
    rk = a.rescore_k if a.precision != "fp32" else 0
    sc = 0 if (a.precision == "fp32" or rk == 0 or a.search_components == "0") else (
        "auto" if a.search_components == "auto" else int(a.search_components))
    eng, build_ms = None, []
    for rep in range(2):
        eng = None
        torch.cuda.empty_cache()
        torch.cuda.synchronize()
        t_b = time.perf_counter()
        eng = E.DictionaryIndexer(mp, rays, dico, precision=a.precision, rescore_k=rk, search_components=sc,
                                  explained_variance=a.explained_variance)
        torch.cuda.synchronize()
        build_ms.append((time.perf_counter() - t_b) * 1000)
    n_dict = eng.size
    search_dim = eng.search_components if eng.search_components else D
    search_cfg = ({"pca_components": eng.search_components, "second_moment_fraction": round(eng.explained, 5),
                   "rescore": f"top-{max(rk, a.topk)} exactly in FP32 on the full {D} pixels"}
                  if eng.search_components else f"full {D} pixels")
    pca_search = bool(eng.search_components)
    # the projection kernel alone, on the dtype it writes (with a PCA search:
    # the FP32 dictionary the basis and the rescoring come from)
    proj_dtype = DTYPE_F32 if eng.search_components else eng.dtype
    proj_out = eng.dictionary_f32 if eng.search_components else eng.dictionary
    pe = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    pe[0].record()
    kernels.project(mp, rays, dico, proj_dtype, out=proj_out, quads=eng.quads)
    pe[1].record()
    torch.cuda.synchronize()
    proj_kernel_ms = pe[0].elapsed_time(pe[1])
    log(f"dictionary build: {build_ms[-1]:.1f} ms (engine), projection kernel {proj_kernel_ms:.2f} ms"
        + (f"; search on {eng.search_components} principal components ({eng.explained:.4f} of the second moment)"
           if eng.search_components else ""))

    head_kernel = sim_kernel_label(eng, a.batch, max(a.topk, rk))
    head = time_form(eng, exp, a.topk, a.steps, a.warmup)
    value = head["rate"]
    log(f"{a.precision}: {head['ms_per_step']:.1f} ms/step, similarity {head['sim_ms']:.1f} ms, {value:.0f} patterns/s")

    flops = 2.0 * a.batch * n_dict * search_dim
    achieved = flops / (head["sim_ms"] / 1000) / 1e12
    peak = PEAKS[a.precision]
    out_bytes = 4 if pca_search else {"fp32": 4, "bf16": 2, "fp16": 2, "fp8": 1}[a.precision]
    proj_bytes = n_dict * (D * out_bytes + 16) + 2 * a.mp_size * a.mp_size * 4 + 12 * D
    proj_gbps = proj_bytes / (proj_kernel_ms / 1000) / 1e9

    # ---- plain FP32 DI over the whole batch: agreement, and its own rate ----
    agree = head_mis = fp32 = None
    ho_n, ho_agree = min(16384, a.batch), None
    head_cos = fp32_cos = None
    forms = {}
    if a.agree:
        ref = E.DictionaryIndexer(mp, rays, dico, precision="fp32")
        fp32 = time_form(ref, exp, a.topk, 1, 0, warmup_rows=2048)
        q32 = ref.prepare(exp)
        agree = float((fp32["idx"][:, 0] == head["idx"][:, 0]).float().mean())
        head_mis = mismatch_report(q32, ref.dictionary, head["idx"][:, 0], fp32["idx"][:, 0])
        log(f"top-1 agreement with FP32 over all {a.batch} patterns: {agree:.5f}; FP32 DI {fp32['rate']:.0f} "
            "patterns/s")
        # held out: orientations and noise the component count was not chosen on
        ho_ori = synthetic.random_orientations(ho_n, seed=202, device=dev)
        ho_exp = synthetic.simulate_patterns(mp, rays, ho_ori, mean_counts=200.0, gaussian_sigma=0.05, seed=203)
        ho_agree = float((form_top1(ref, ho_exp) == form_top1(eng, ho_exp)).float().mean())
        log(f"held-out top-1 agreement ({ho_n} patterns, orientation seed 202, noise seed 203): {ho_agree:.5f}")
        head_cos = {"reported_top1": cos_error(q32, ref.dictionary, head["idx"][:, 0], head["val"][:, 0])}
        if head["search_idx"] is not None:
            head_cos["search_stage_top1"] = cos_error(q32, ref.dictionary, head["search_idx"][:, 0],
                                                      head["search_val"][:, 0])
        fp32_cos = {"reported_top1": cos_error(q32, ref.dictionary, fp32["idx"][:, 0], fp32["val"][:, 0]),
                    "search_stage_top1": cos_error(q32, ref.dictionary, fp32["search_idx"][:, 0],
                                                   fp32["search_val"][:, 0])}
        if a.forms and a.precision != "fp32":
            del eng
            torch.cuda.empty_cache()
            for name, prec, comps, ev in (("bf16_full_patterns", "bf16", 0, None),
                                          ("fp8_full_patterns", "fp8", 0, None),
                                          ("fp8_pca_library_default", "fp8", "auto", 0.995)):
                forms[name] = sub_form(mp, rays, dico, exp, ho_exp, ref, q32, fp32, prec, comps, ev, a)
                torch.cuda.empty_cache()
            eng = None
        del ref, q32
        torch.cuda.empty_cache()

    refinement = None
    if a.refine:
        refinement = refinement_form(a, mp, geom, exp, ori, dico[head["idx"][:, 0].long()], det, dev)

    nlpar_res = None
    if a.nlpar:
        nlpar_res = nlpar_form(a, exp, det, dev)

    pcfit = None
    if a.pc_fit:
        pcfit = pc_fit_form(a, mp, exp, det, dev)

    learned = None
    if a.learned_train > 0:
        eng = None
        torch.cuda.empty_cache()
        learned = learned_form(a, mp, dico, dev)

    # HBM-side bytes per launch of this kernel at this workload, measured
    # with two PMC passes (FETCH_SIZE, WRITE_SIZE) and committed under
    # profiles/ (null when the workload differs from the measured one)
    traffic = lookup_traffic(a.precision, n_dict, a.batch, [det, det], a.topk, search_dim, D)

    cpu = c1 = None
    if a.cpu_seconds > 0:
        cpu = cpu_baseline(mp.cpu(), rays.cpu(), exp[:1024].cpu(), dico.cpu(), n_dict, a.cpu_seconds)
        c1 = c1_form(a, rays, dev)

    build_s = build_ms[-1] / 1000
    res = {
        "metric": "patterns indexed/s vs plain FP32 DI at >=99.9% top-1 match (1M dict, 60x60)",
        "value": round(value, 1),
        "unit": "patterns/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
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
        "top1_agreement_vs_fp32": agree,
        "agreement_sample": None if agree is None else a.batch,
        "top1_disagreements_vs_fp32": head_mis,
        "heldout_top1_agreement_vs_fp32": {"value": ho_agree, "patterns": ho_n,
                                           "seeds": "orientations 202, noise 203 (the batch: 2, 3)"},
        "fp32_di_patterns_per_s": None if fp32 is None else round(fp32["rate"], 1),
        "speedup_vs_fp32_di": None if fp32 is None else round(value / fp32["rate"], 2),
        "including_dictionary_build": {
            "build_s": round(build_s, 3),
            "patterns_per_s_one_batch": round(a.batch / (build_s + head["ms_per_step"] / 1000), 1),
            "what": "one batch indexed right after building the dictionary engine (projection, second moment, "
                    "FP64 leading eigenvectors, coordinates); the build is once per dictionary",
        },
        "cosine_error_vs_fp64": None if head_cos is None else {
            a.precision + ("_pca" if pca_search else "") + "+rescore": head_cos,
            "fp32": fp32_cos,
            "what": "|s - <q,d>_fp64| / (|q| |d|) for the reported top-1 of every pattern, <q,d>_fp64 the exact "
                    "dot product of the same FP32 zero-mean rows; search_stage_top1 is the low-precision "
                    "search's own score of its top candidate (for a principal-coordinate search it also "
                    "holds the truncation to the leading components)",
        },
        "forms": forms,
        "roofline": {
            "kernel": "sim_topk (similarity GEMM + fused top-k) " + head_kernel,
            "bound": "mfma",
            "achieved": round(achieved, 2),
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": traffic,
            "avg_launch_ms": round(head["sim_ms"], 3),
            "flops_per_launch": flops,
            "contraction_dim": search_dim,
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
        "c1": c1,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


def time_form(eng, exp, topk, steps, warmup, warmup_rows=None):
    """Index `exp` with a DictionaryIndexer: `warmup` untimed passes (on
    the first warmup_rows patterns if given), then `steps` timed passes of
    preparation + similarity/top-k (+ FP32 rescore). Returns the wall rate,
    ms per step, the similarity kernel's average launch time (HIP events on
    torch's current stream, where the library launches) and the last pass's
    results (and the search stage's own list when there is a rescore)."""
    import torch

    from ebsdtorch_amd import kernels

    dev = exp.device
    sims = []

    def one(x):
        q, qf = eng.prepare_pair(x)
        has_f32 = eng.dictionary_f32 is not None
        k_search = max(topk, eng.rescore_k) if has_f32 else topk
        val = torch.empty((q.rows, k_search), dtype=torch.float32, device=dev)
        idx = torch.empty((q.rows, k_search), dtype=torch.int32, device=dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        kernels.sim_topk(q, eng.dictionary, k_search, val, idx, merge=False)
        e1.record()
        sims.append((e0, e1))
        if has_f32:
            v2, i2 = kernels.rescore(qf, eng.dictionary_f32, idx, topk)
            return v2, i2, val, idx
        return val, idx, None, None

    if warmup_rows:
        one(exp[:warmup_rows])
    for _ in range(warmup):
        one(exp)
    torch.cuda.synchronize()
    sims.clear()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = one(exp)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sim_ms = sum(e0.elapsed_time(e1) for e0, e1 in sims) / len(sims)
    return {"rate": exp.shape[0] * steps / wall, "ms_per_step": wall * 1000 / steps, "sim_ms": sim_ms,
            "val": out[0], "idx": out[1], "search_val": out[2], "search_idx": out[3]}


def mismatch_report(q32, d32, idx_a, idx_b, tie=1e-6):
    """Top-1 disagreements of two forms: how many, and how many are not
    ties, i.e. whose exact (FP64) scores of the two candidates differ by at
    least `tie` in cosine units (|s_a - s_b| / (|q| max(|d_a|, |d_b|)))."""
    import torch

    diff = torch.nonzero(idx_a != idx_b).flatten()
    if diff.numel() == 0:
        return {"mismatches": 0, "beyond_ties": 0}
    q = q32.data[diff, :q32.n_pix].double()
    da = d32.data[idx_a[diff].long(), :q32.n_pix].double()
    db = d32.data[idx_b[diff].long(), :q32.n_pix].double()
    gap = ((q * da).sum(1) - (q * db).sum(1)).abs() / (q.norm(dim=1) * torch.maximum(da.norm(dim=1), db.norm(dim=1)))
    return {"mismatches": int(diff.numel()), "beyond_ties": int((gap >= tie).sum()),
            "max_gap_cosine_units": float(f"{float(gap.max()):.3g}")}


def form_top1(eng, x):
    """Top-1 dictionary index of every row of x (int32, on the device)."""
    q, qf = eng.prepare_pair(x)
    return eng.index(q, 1, queries_f32=qf)[1][:, 0]


def cos_error(q32, d32, idx, score, chunk=8192):
    """Quantiles (p50, p99, p99.9, max) over the rows of
    |score - <q, d_idx>| / (|q| |d_idx|) with the dot product and the norms
    taken in float64 from the FP32 rows of q32 (queries) and d32 (the FP32
    dictionary)."""
    import torch

    n_pix = q32.n_pix
    errs = []
    for s in range(0, q32.rows, chunk):
        q = q32.data[s:s + chunk, :n_pix].double()
        d = d32.data[idx[s:s + chunk].long(), :n_pix].double()
        exact = (q * d).sum(1)
        errs.append((score[s:s + chunk].double() - exact).abs() / (q.norm(dim=1) * d.norm(dim=1)))
    e = torch.cat(errs)
    qs = torch.tensor([0.5, 0.99, 0.999], dtype=torch.float64, device=e.device)
    p = torch.quantile(e, qs).tolist()
    return {"p50": float(f"{p[0]:.3g}"), "p99": float(f"{p[1]:.3g}"), "p99.9": float(f"{p[2]:.3g}"),
            "max": float(f"{float(e.max()):.3g}")}


def sub_form(mp, rays, dico, exp, ho_exp, ref, q32, fp32, precision, components, ev, a):
    """One more precision form of the metric workload, timed over the batch
    (a few steps): patterns/s, similarity kernel time and MFMA fraction, top-1
    agreement with plain FP32 DI over the whole batch and on the held-out
    set, cosine error against FP64."""
    import torch

    import ebsdtorch_amd as E

    torch.cuda.synchronize()
    t_b = time.perf_counter()
    kw = {} if ev is None else {"explained_variance": ev}
    eng = E.DictionaryIndexer(mp, rays, dico, precision=precision, rescore_k=16, search_components=components,
                              **kw)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t_b
    steps = max(1, min(a.steps, 3))
    r = time_form(eng, exp, a.topk, steps, 1)
    k = eng.search_components or exp.shape[1]
    flops = 2.0 * exp.shape[0] * eng.size * k
    tf = flops / (r["sim_ms"] / 1000) / 1e12
    out = {
        "precision": precision,
        "search": (f"{eng.search_components} principal components ({eng.explained:.4f} of the second moment)"
                   if eng.search_components else f"full {exp.shape[1]} pixels"),
        "patterns_per_s": round(r["rate"], 1), "ms_per_step": round(r["ms_per_step"], 2), "steps": steps,
        "similarity_kernel": {"name": sim_kernel_label(eng, exp.shape[0], max(a.topk, 16)),
                              "avg_launch_ms": round(r["sim_ms"], 3), "achieved_TFLOPs": round(tf, 1),
                              "peak_TFLOPs": PEAKS[precision], "mfma_frac": round(tf / PEAKS[precision], 4),
                              "contraction_dim": k},
        "top1_agreement_vs_fp32": float((r["idx"][:, 0] == fp32["idx"][:, 0]).float().mean()),
        "top1_disagreements_vs_fp32": mismatch_report(q32, ref.dictionary, r["idx"][:, 0], fp32["idx"][:, 0]),
        "heldout_top1_agreement_vs_fp32": float((form_top1(eng, ho_exp) == form_top1(ref, ho_exp)).float().mean()),
        "cosine_error_vs_fp64": {
            "reported_top1": cos_error(q32, ref.dictionary, r["idx"][:, 0], r["val"][:, 0]),
            "search_stage_top1": cos_error(q32, ref.dictionary, r["search_idx"][:, 0], r["search_val"][:, 0])},
        "build_s": round(build_s, 3),
    }
    log(f"form {precision} ({out['search']}): {out['patterns_per_s']:.0f} patterns/s, similarity "
        f"{r['sim_ms']:.1f} ms ({out['similarity_kernel']['mfma_frac']:.3f} of peak), agreement "
        f"{out['top1_agreement_vs_fp32']:.5f} / held out {out['heldout_top1_agreement_vs_fp32']:.5f}")
    del eng
    return out


def sim_kernel_label(eng, m, topk):
    """Name and tile of the similarity kernel the library launches for m
    queries against the engine's search matrix (from ebsd_sim_schedule)."""
    from ebsdtorch_amd import kernels

    d = eng.dictionary
    plan = kernels.sim_schedule(m, d.rows, topk, d.dtype, d.ld, counts=False)
    name = {2: "sim_topk_b2_kernel", 1: "sim_topk_bt_kernel", 0: "sim_topk_kernel"}[plan["bt"]]
    return f"{name} {plan['d_rows']}x{plan['q_rows']}"


def lookup_traffic(precision, n_dict, batch, detector, topk, search_dim, D):
    """HBM-side bytes per similarity launch recorded from PMC passes in
    profiles/sim_topk_traffic.json for this exact workload, else None."""
    tpath = os.path.join(ROOT, "profiles", "sim_topk_traffic.json")
    if not os.path.exists(tpath):
        return None
    with open(tpath) as f:
        recs = json.load(f)
    for t in (recs if isinstance(recs, list) else [recs]):
        if (t.get("precision") == precision and t.get("dictionary") == n_dict and t.get("batch") == batch
                and t.get("detector") == detector and t.get("topk") == topk
                and t.get("search_dim", D) == search_dim):
            return t["traffic_bytes"]
    return None



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
    # the threshold keeps 99.99% of the accepted calibration patterns within
    # 1 degree, on 32,768 of them. The rare wrong proposals (a wrong basin,
    # tens of degrees off) score below ~0.6 while almost every right one
    # scores above 0.7, so the strict target costs ~0.1% more fallback; a
    # 99.9% target on 8,192 patterns allows ~8 misses and put the threshold
    # anywhere in 0.08-0.38 from one seed to the next, letting the held-out
    # batch fall below 99.9% within 1 degree
    idx.threshold = calibrate_threshold(idx, cal, n=32768, max_error_deg=1.0, precision=0.9999)
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


def projection_workload(a, dev):
    """The projection kernel alone (for counters): the metric dictionary
    (995,277 m-3m orientations, 60x60), the same at 80x80 (C3) or the C5 one
    (986,945 6/mmm orientations, 120x120, 60 band families), projected with zero mean into
    FP32 rows and into FP8 rows (per-row power-of-two scale), `steps` times
    each after one warm-up; HIP-event times and algorithmic GB/s (output rows
    + 16 bytes of quaternion per orientation)."""
    import torch

    import ebsdtorch_amd as E
    from ebsdtorch_amd import kernels, synthetic
    from ebsdtorch_amd._lib import DTYPE_F32, DTYPE_FP8_E4M3

    laue, det, fam, seed = {"metric": (11, 60, 40, 0), "c3": (11, 80, 40, 0), "c5": (9, 120, 60, 10)}[a.proj_config]
    mp = synthetic.band_master_pattern(a.mp_size, laue, fam, 2.0, seed=seed, device=dev)
    dico = E.sample_ori_fz_laue(laue, a.dict_size, dev, permute=True,
                                generator=torch.Generator().manual_seed(1 if laue == 11 else 11))
    rays = E.EBSDGeometry((det, det), proj_center=(0.5, 0.5, 0.6)).get_rays().to(dev)
    quads = kernels.mp_quad_table(mp)
    D = det * det
    out = {}
    if a.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        torch.empty_like(src).copy_(src)
        torch.cuda.synchronize()
        del src
    for name, dt, eb in (("fp32", DTYPE_F32, 4), ("fp8", DTYPE_FP8_E4M3, 1)):
        o = kernels.alloc_patterns(dico.shape[0], D, dt, dev)
        kernels.project(mp, rays, dico, dt, out=o, quads=quads)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(a.steps):
            kernels.project(mp, rays, dico, dt, out=o, quads=quads)
        ev[1].record()
        torch.cuda.synchronize()
        ms = ev[0].elapsed_time(ev[1]) / a.steps
        nbytes = dico.shape[0] * (o.ld * eb + (4 if dt == DTYPE_FP8_E4M3 else 0) + 16)
        out[name] = {"ms": round(ms, 3), "GBps": round(nbytes / ms / 1e6, 1),
                     "hbm_frac": round(nbytes / ms / 1e6 / PEAKS["hbm_GBps"], 4), "bytes": nbytes,
                     "pixels_per_s": round(dico.shape[0] * D / ms * 1000, 1)}
        log(f"projection {a.proj_config} {name}: {ms:.2f} ms, {out[name]['GBps']} GB/s")
        del o
    return {"workload": f"projection of {dico.shape[0]} orientations onto {det}x{det} "
                        f"({'m-3m' if laue == 11 else '6/mmm'} {a.mp_size}x{a.mp_size} master pattern)",
            "steps": a.steps, **out}



def library_gemm_workload(a, dev):
    """The similarity contraction as a plain library GEMM (hipBLASLt through
    torch): 65,536 query rows x 16,384 dictionary rows per call at the row
    widths the search kernels use (FP8 3,712 and 256 bytes, BF16 3,648
    elements), BF16 output (2.1 GB per call, written; no top-k). A
    reference point for the fused kernels, which compute the same products
    and keep only the top 16 per query. HIP-event time per call over
    `steps` calls after one warm-up."""
    import torch

    m, n = a.batch, 16384
    gen = torch.Generator(device=dev).manual_seed(0)
    out = {}
    cases = [("fp8", 3712), ("fp8", 256), ("bf16", 3648)]
    for name, k in cases:
        x = torch.randn(m, k, device=dev, generator=gen)
        y = torch.randn(n, k, device=dev, generator=gen)
        try:
            if name == "fp8":
                x8, y8 = x.to(torch.float8_e4m3fn), y.to(torch.float8_e4m3fn)
                one = torch.ones((), device=dev)

                def run():
                    return torch._scaled_mm(x8, y8.t(), scale_a=one, scale_b=one, out_dtype=torch.bfloat16)
            else:
                xb, yb = x.to(torch.bfloat16), y.to(torch.bfloat16)

                def run():
                    return torch.matmul(xb, yb.t())
            run()
            torch.cuda.synchronize()
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record()
            for _ in range(a.steps):
                run()
            ev[1].record()
            torch.cuda.synchronize()
            ms = ev[0].elapsed_time(ev[1]) / a.steps
            tflops = 2.0 * m * n * k / ms / 1e9
            out[f"{name}_k{k}"] = {"ms": round(ms, 3), "TFLOPs": round(tflops, 1),
                                   "frac": round(tflops / PEAKS[name], 4)}
            log(f"library GEMM {name} {m}x{n}x{k}: {ms:.3f} ms, {tflops:.1f} TFLOP/s")
        except Exception as e:  # report an unsupported case instead of failing the run
            out[f"{name}_k{k}"] = {"error": f"{type(e).__name__}: {e}"}
            log(f"library GEMM {name} k={k}: {type(e).__name__}: {e}")
        torch.cuda.empty_cache()
    return {"workload": f"library GEMM (hipBLASLt via torch) {m} x {n} per call, BF16 output", "steps": a.steps,
            **out}


def c3_workload(a, dev):
    """The precision configuration: 995,277-entry m-3m dictionary projected
    onto an 80x80 detector (D = 6,400) from the 501x501 band master pattern,
    65,536 noisy patterns (Poisson 200 + N(0, 0.05^2)), top-16. Forms, each
    over the whole batch: plain FP32 DI (the reference for agreement), BF16
    and FP8 on the full patterns and FP8 on the leading principal
    coordinates, each with the top 16 rescored exactly in FP32; and PCA DI
    (indexing.py:214-476) with 1,000 components, covariance accumulated in
    FP32 on the matrix cores over the first `c3_learn` dictionary rows, the
    eigensolve in FP64, the search in FP32 on the coefficients. Per form:
    patterns/s, the similarity kernel's MFMA fraction, top-1 agreement with
    FP32 DI, and the cosine error of the reported top-1 against FP64."""
    import torch

    import ebsdtorch_amd as E
    from ebsdtorch_amd import kernels, synthetic
    from ebsdtorch_amd._lib import DTYPE_F32

    det, topk = 80, 16
    D = det * det
    t0 = time.perf_counter()
    mp = synthetic.band_master_pattern(a.mp_size, 11, 40, 2.0, seed=0, device=dev)
    dico = E.sample_ori_fz_laue(11, a.dict_size, dev, permute=True, generator=torch.Generator().manual_seed(1))
    geom = E.EBSDGeometry((det, det), proj_center=(0.5, 0.5, 0.6))
    rays = geom.get_rays().to(dev)
    ori = synthetic.random_orientations(a.batch, seed=2, device=dev)
    exp = synthetic.simulate_patterns(mp, rays, ori, mean_counts=200.0, gaussian_sigma=0.05, seed=3)
    torch.cuda.synchronize()
    n_dict = dico.shape[0]
    log(f"c3: inputs ready in {time.perf_counter() - t0:.1f}s: dictionary {n_dict}, batch {a.batch}, D={D}")
    forms = {}

    # plain FP32 DI: the agreement reference (one pass over the batch)
    torch.cuda.synchronize()
    tb = time.perf_counter()
    ref = E.DictionaryIndexer(mp, rays, dico, precision="fp32")
    torch.cuda.synchronize()
    build32 = time.perf_counter() - tb
    pe = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    pe[0].record()
    kernels.project(mp, rays, dico, DTYPE_F32, out=ref.dictionary, quads=ref.quads)
    pe[1].record()
    torch.cuda.synchronize()
    proj_ms = pe[0].elapsed_time(pe[1])
    proj_bytes = n_dict * (ref.dictionary.ld * 4 + 16) + 2 * a.mp_size * a.mp_size * 4 + 12 * D
    fp32 = time_form(ref, exp, topk, 1, 0, warmup_rows=2048)
    q32 = ref.prepare(exp)
    flops32 = 2.0 * a.batch * n_dict * D
    forms["fp32"] = {
        "patterns_per_s": round(fp32["rate"], 1), "build_s": round(build32, 3),
        "similarity_kernel": {"avg_launch_ms": round(fp32["sim_ms"], 2),
                              "mfma_frac": round(flops32 / (fp32["sim_ms"] / 1000) / 1e12 / PEAKS["fp32"], 4)},
        "top1_agreement_vs_fp32": 1.0,
        "cosine_error_vs_fp64": {
            "reported_top1": cos_error(q32, ref.dictionary, fp32["idx"][:, 0], fp32["val"][:, 0]),
            "search_stage_top1": cos_error(q32, ref.dictionary, fp32["search_idx"][:, 0],
                                           fp32["search_val"][:, 0])},
    }
    log(f"c3 fp32: {fp32['rate']:.0f} patterns/s, similarity {fp32['sim_ms']:.0f} ms")
    ho_ori = synthetic.random_orientations(8192, seed=202, device=dev)
    ho_exp = synthetic.simulate_patterns(mp, rays, ho_ori, mean_counts=200.0, gaussian_sigma=0.05, seed=203)
    for name, prec, comps, ev in (("bf16_full_patterns", "bf16", 0, None),
                                  ("fp8_full_patterns", "fp8", 0, None),
                                  ("fp8_pca", "fp8", "auto", a.explained_variance)):
        forms[name] = sub_form(mp, rays, dico, exp, ho_exp, ref, q32, fp32, prec, comps, ev, a)
        torch.cuda.empty_cache()

    # PCA DI (the reference's second indexing method) with 1,000 components
    n_learn = min(a.c3_learn, n_dict)
    torch.cuda.synchronize()
    tb = time.perf_counter()
    coords = geom.get_coords_sample_frame((1, 1)).view(-1, 3).to(dev)
    pca = E.PCADictionaryIndexer(mp, coords, dico[:n_learn], a.c3_pca_components, index_orientations=dico)
    torch.cuda.synchronize()
    build_pca = time.perf_counter() - tb
    tm = pca.timings
    cov_flops = 1.0 * n_learn * D * (D + 1)  # the symmetric update: upper triangle with the diagonal
    coef_flops = 2.0 * n_dict * D * pca.n_components
    pca_r = time_form(pca, exp, topk, 2, 1)
    s_flops = 2.0 * a.batch * n_dict * pca.n_components
    # cosine error of the PCA-space score against the FP64 PCA-space dot
    # product (its arithmetic), and against the full-pattern exact dot
    # product (arithmetic + truncation to the components)
    comp = pca.components  # (D, n) float64
    top1 = pca_r["idx"][:, 0].long()
    errs = []
    for s0 in range(0, a.batch, 4096):
        qz = q32.data[s0:s0 + 4096, :D].double() @ comp
        dz = ref.dictionary.data[top1[s0:s0 + 4096], :D].double() @ comp
        errs.append((pca_r["val"][s0:s0 + 4096, 0].double() - (qz * dz).sum(1)).abs()
                    / (qz.norm(dim=1) * dz.norm(dim=1)))
    e = torch.cat(errs)
    qs = torch.tensor([0.5, 0.99, 0.999], dtype=torch.float64, device=e.device)
    pq = torch.quantile(e, qs).tolist()
    forms["pca_di"] = {
        "components": pca.n_components, "learning_rows": n_learn,
        "explained": round(float(pca.eigenvalues[-pca.n_components:].sum() / pca.eigenvalues.sum()), 5),
        "patterns_per_s": round(pca_r["rate"], 1), "build_s": round(build_pca, 2),
        "build_stages_ms": {k: round(v, 1) for k, v in tm.items()},
        "covariance": {"ms": round(tm["covariance_ms"], 1), "flops": cov_flops,
                       "mfma_frac_fp32": round(cov_flops / (tm["covariance_ms"] / 1000) / 1e12 / PEAKS["fp32"], 4)},
        "coefficients_gemm": {"ms": round(tm["index_coefficients_ms"], 1), "flops": coef_flops,
                              "mfma_frac_fp32": round(coef_flops / (tm["index_coefficients_ms"] / 1000) / 1e12
                                                      / PEAKS["fp32"], 4)},
        "eigensolve": "FP64, host LAPACK, " + f"{tm['eigensolve_ms'] / 1000:.1f} s",
        "similarity_kernel": {"avg_launch_ms": round(pca_r["sim_ms"], 2),
                              "mfma_frac": round(s_flops / (pca_r["sim_ms"] / 1000) / 1e12 / PEAKS["fp32"], 4),
                              "contraction_dim": pca.n_components},
        "top1_agreement_vs_fp32": float((pca_r["idx"][:, 0] == fp32["idx"][:, 0]).float().mean()),
        "top1_disagreements_vs_fp32": mismatch_report(q32, ref.dictionary, pca_r["idx"][:, 0], fp32["idx"][:, 0]),
        "cosine_error_vs_fp64": {
            "pca_space": {"p50": float(f"{pq[0]:.3g}"), "p99": float(f"{pq[1]:.3g}"),
                          "p99.9": float(f"{pq[2]:.3g}"), "max": float(f"{float(e.max()):.3g}")},
            "vs_full_pattern_dot": cos_error(q32, ref.dictionary, pca_r["idx"][:, 0], pca_r["val"][:, 0])},
    }
    log(f"c3 pca di: {pca_r['rate']:.0f} patterns/s, build {build_pca:.1f} s ({tm}), agreement "
        f"{forms['pca_di']['top1_agreement_vs_fp32']:.4f}")
    return {
        "metric": "c3: patterns indexed/s, 1M m-3m dictionary, 80x80, per precision form",
        "workload": f"DI {n_dict} m-3m dictionary x {a.batch}-pattern batch, {det}x{det} detector, top-{topk}",
        "n_gpus": 1,
        "data": "synthetic: 501x501 m-3m band master pattern (40 families, 2 deg), noisy patterns "
                "(Poisson 200 + N(0,0.05^2)) at seeded uniform orientations; held-out set: seeds 202 / 203",
        "projection_fp32": {"ms": round(proj_ms, 2), "GBps": round(proj_bytes / proj_ms / 1e6, 1),
                            "hbm_frac": round(proj_bytes / proj_ms / 1e6 / PEAKS["hbm_GBps"], 4)},
        "forms": forms,
    }


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
        n=8192, precision=0.999)
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


def c1_form(a, rays, dev):
    """The small configuration in full on both sides: 201x201 m-3m band
    master pattern, the ~10k-entry dictionary (sample_ori_fz_laue(11,
    10,000)), 1,000 noisy 60x60 patterns, top-1. GPU: plain FP32 DI
    (DictionaryIndexer), with and without its dictionary build. CPU: the
    oracle's dictionary_index (projection + FP32 matching at the
    reference's 256 x 256 chunking) on all 1,000 patterns, and the int8
    variant. Top-1 agreement of the two, with the disagreements' gaps."""
    import torch

    import ebsdtorch_amd as E
    from ebsdtorch_amd import synthetic
    from oracle import di_oracle as O

    mp = synthetic.band_master_pattern(201, 11, 40, 2.0, seed=0, device=dev)
    dico = E.sample_ori_fz_laue(11, 10_000, dev, permute=True, generator=torch.Generator().manual_seed(1))
    ori = synthetic.random_orientations(1000, seed=2, device=dev)
    exp = synthetic.simulate_patterns(mp, rays, ori, mean_counts=200.0, gaussian_sigma=0.05, seed=3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng = E.DictionaryIndexer(mp, rays, dico, precision="fp32")
    gidx = form_top1(eng, exp)
    torch.cuda.synchronize()
    with_build = time.perf_counter() - t0
    run = time_form(eng, exp, 1, 5, 1)
    mp_c, rays_c, exp_c, dico_c = mp.cpu(), rays.cpu(), exp.cpu(), dico.cpu()
    t0 = time.perf_counter()
    cidx, _ = O.dictionary_index(mp_c, rays_c, exp_c, dico_c, top_k=1)
    t_cpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.dictionary_index(mp_c, rays_c, exp_c, dico_c, top_k=1, quantized=True)
    t_int8 = time.perf_counter() - t0
    q32 = eng.prepare(exp)
    rep = mismatch_report(q32, eng.dictionary, gidx, cidx[:, 0].to(dev).int())
    agree = float((gidx.cpu().long() == cidx[:, 0]).float().mean())
    log(f"C1: GPU FP32 DI {run['rate']:.0f} patterns/s ({1000 / with_build:.0f} with the build), CPU oracle "
        f"{1000 / t_cpu:.1f} patterns/s (int8 {1000 / t_int8:.1f}), top-1 agreement {agree:.4f}")
    return {"workload": f"C1: {dico.shape[0]} m-3m dictionary x 1000 patterns, 60x60, 201x201 master pattern, top-1",
            "gpu_fp32_patterns_per_s": round(run["rate"], 1),
            "gpu_fp32_patterns_per_s_with_build": round(1000 / with_build, 1),
            "cpu_oracle_patterns_per_s": round(1000 / t_cpu, 2),
            "cpu_oracle_int8_patterns_per_s": round(1000 / t_int8, 2),
            "cpu_cores": torch.get_num_threads(),
            "top1_agreement_gpu_vs_cpu_oracle": agree, "disagreements": rep,
            "what": "the CPU figures include the oracle's own dictionary projection (as the reference's call "
                    "does); the first GPU figure excludes the one-off dictionary build, the second includes it"}


def cpu_baseline(mp, rays, exp, dico, n_dict, seconds):
    """The oracle on the host cores, two variants of the reference's CPU DI,
    each on 1,024 patterns against a dictionary slice sized for ~`seconds`
    of work, scaled to the full dictionary: plain FP32 (the oracle's
    semantics) at the reference's default chunking (256 dictionary x 256
    experimental patterns, indexing.py:20-22) and at its demo's (4096 x 4096,
    indexing.py:533-536), `value` being the better; and the reference's actual
    CPU default, int8 dynamic quantisation of each dictionary chunk
    (quantized_if_x86, knn.py:52-72; `int8`, 4096 chunking)."""
    import platform
    import warnings

    import torch

    from oracle import di_oracle as O

    warnings.filterwarnings("ignore", message=".*torch.ao.quantization is deprecated.*")
    cores = torch.get_num_threads()

    def timed(quantized, budget, chunk):
        n_sub = 4096
        while True:
            t0 = time.perf_counter()
            O.dictionary_index(mp, rays, exp, dico[:n_sub], top_k=16, dictionary_chunk_size=chunk,
                               experiment_chunk_size=chunk, quantized=quantized)
            t = time.perf_counter() - t0
            if t >= budget / 3 or n_sub >= n_dict:
                return n_sub, t
            n_sub = min(n_dict, int(n_sub * max(2.0, budget / max(t, 1e-3))))

    runs = {}
    for chunk, budget in ((4096, seconds * 0.5), (256, seconds * 0.5)):
        n_sub, t = timed(False, budget, chunk)
        runs[chunk] = (exp.shape[0] / (t * n_dict / n_sub), n_sub, t)
        log(f"cpu baseline {chunk}x{chunk}: {exp.shape[0]} patterns x {n_sub} dictionary entries in {t:.1f}s "
            f"on {cores} threads")
    n8, t8 = timed(True, seconds / 2, 4096)
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
    best = max(runs, key=lambda c: runs[c][0])
    rate, n_sub, t = runs[best]
    return {"value": round(rate, 3), "unit": "patterns/s", "cores": cores, "kind": "port",
            "sample": f"{exp.shape[0]} patterns x first {n_sub} of {n_dict} dictionary entries "
                      f"({t:.1f} s), scaled to the full dictionary; chunking {best} x {best}",
            "by_chunking": {f"{c}x{c}": {"value": round(r, 3), "dictionary_entries": n, "seconds": round(tt, 2)}
                            for c, (r, n, tt) in runs.items()},
            "int8": {"value": round(rate8, 3), "unit": "patterns/s",
                     "what": "the reference's CPU default: int8 quantize_dynamic of each dictionary chunk",
                     "sample": f"{exp.shape[0]} patterns x first {n8} dictionary entries ({t8:.1f} s), scaled"},
            "host": {"os_cpu_count": os.cpu_count(), "threads_used": cores, "model": model, "libraries": libs}}


if __name__ == "__main__":
    main()
