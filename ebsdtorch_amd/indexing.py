"""Dictionary indexing drivers (ebsd/indexing.py:14-211 and :214-476).

dictionary_index_orientations keeps the reference's signature and side
effects (results written into the ExperimentPatterns object). Underneath,
the dictionary is projected by the fused HIP projection straight into the
layout the similarity kernel reads, and every query is scored against it by
the MFMA similarity kernel with the top-k fused in; the reference's chunk
sizes are accepted but the engine picks its own (large) chunks, which cannot
change the result because the top-k order is total (score, then index).

DictionaryIndexer is the reusable engine behind the driver: it holds one
projected dictionary resident in HBM and indexes batches of patterns
against it.
"""

from __future__ import annotations

import time
from typing import List, Optional, Sequence, Union

import torch

from . import kernels
from ._lib import DTYPE_BF16, DTYPE_F32, DTYPE_FP8_E4M3, EBSD_MAX_TOPK
from .experiment_pats import ExperimentPatterns
from .geometry import EBSDGeometry
from .knn import ChunkedKNN
from .master_pattern import MasterPattern
from .orientations import sample_ori_fz_laue_angle


def _gpu_device(t: torch.Tensor) -> torch.device:
    if t.device.type == "cuda":
        return t.device
    if not torch.cuda.is_available():
        raise RuntimeError("dictionary indexing runs on the GPU and no HIP device is visible")
    return torch.device("cuda", torch.cuda.current_device())


class _ResidentSearch:
    """Top-k search of prepared queries against a resident pattern matrix
    (self.dictionary in the search dtype, optional FP32 copy for rescoring)."""

    dictionary: kernels.PatternMatrix
    dictionary_f32: Optional[kernels.PatternMatrix]
    rescore_k: int
    device: torch.device

    def index(self, queries: kernels.PatternMatrix, topk: int, queries_f32: Optional[kernels.PatternMatrix] = None):
        """Top-k (values (N, k) f32, indices (N, k) int32) of prepared queries."""
        m = queries.rows
        k_search = max(topk, self.rescore_k) if self.dictionary_f32 is not None else topk
        val = torch.empty((m, k_search), dtype=torch.float32, device=self.device)
        idx = torch.empty((m, k_search), dtype=torch.int32, device=self.device)
        kernels.sim_topk(queries, self.dictionary, k_search, val, idx, index_base=0, merge=False)
        if self.dictionary_f32 is not None:
            if queries_f32 is None and self.dictionary_f32 is self.dictionary:
                queries_f32 = queries  # FP32 search: the queries are their own FP32 rows
            if queries_f32 is None:
                raise ValueError("FP32 rescoring needs the FP32 queries")
            val, idx = kernels.rescore(queries_f32, self.dictionary_f32, idx, topk)
        return val, idx


class DictionaryIndexer(_ResidentSearch):
    """A projected orientation dictionary resident on one GPU.

    master_pattern (S, 2S) f32, rays (D, 3) f32 unit sample-frame
    directions, orientations (N, 4) f32. precision selects the similarity
    arithmetic: "fp32" (the drop-in), "bf16", "fp16" or "fp8"; with
    rescore_k > 0 the low-precision candidates are re-scored in FP32 against
    an FP32 copy of the dictionary. dict_unit_norm also divides each
    dictionary pattern by its norm (cosine ranking); the reference ranks by
    the raw dot product of zero-mean patterns (False).

    FP32 search (exact_fp32, default): the matrix-core search accumulates
    each score in one FP32 chain over all D pixels, which leaves errors of
    up to ~5e-6 of |q||d| at D = 3,600 (a CPU sgemm, with its many partial
    sums, stays near 2e-7), so its best EBSD_MAX_TOPK candidates are
    re-scored by the FP32 rescoring kernel (partial sums over lanes, error
    ~1e-7) and re-ranked: the reported scores and order are those of an
    accurate FP32 dot product.
    """

    def __init__(self, master_pattern: torch.Tensor, rays: torch.Tensor, orientations: torch.Tensor,
                 precision: str = "fp32", rescore_k: int = 0, dict_unit_norm: bool = False,
                 renorm_dirs: bool = False, search_components: Union[int, str] = 0,
                 explained_variance: float = 0.995, basis_sample: int = 1 << 18, rotate_basis: bool = True,
                 exact_fp32: bool = True, basis_solver: str = "subspace"):
        if precision not in kernels.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(kernels.PRECISIONS)}")
        if search_components and rescore_k <= 0:
            raise ValueError("a PCA-compressed search needs rescore_k > 0 (exact FP32 rescoring)")
        self.device = orientations.device
        self.mp = master_pattern.to(self.device, torch.float32).contiguous()
        self.rays = rays.to(self.device, torch.float32).contiguous()
        self.orientations = orientations.to(self.device, torch.float32).contiguous()
        self.precision = precision
        self.dtype = kernels.PRECISIONS[precision]
        self.rescore_k = int(rescore_k)
        self.unit_norm = bool(dict_unit_norm)
        self.renorm_dirs = bool(renorm_dirs)
        self.n_pix = self.rays.shape[0]
        # the master pattern's quad table: one 16-byte gather per bilinear
        # sample in the projection (same values bit for bit)
        self.quads = kernels.mp_quad_table(self.mp)
        self.basis_t = None
        self.basis_t_bf16 = None
        self.rotate_basis = bool(rotate_basis)
        if basis_solver not in ("subspace", "full"):
            raise ValueError("basis_solver must be 'subspace' or 'full'")
        self.basis_solver = basis_solver
        self.search_components = 0
        self.explained = 1.0
        if search_components:
            self.dictionary_f32 = self._project(DTYPE_F32)
            self._build_pca_search(search_components, float(explained_variance), int(basis_sample))
            return
        self.dictionary = self._project(self.dtype)
        self.dictionary_f32 = None
        if self.dtype == DTYPE_F32:
            if exact_fp32:
                # the FP32 rows are their own rescoring copy
                self.dictionary_f32 = self.dictionary
                self.rescore_k = EBSD_MAX_TOPK
        elif self.rescore_k > 0:
            self.dictionary_f32 = self._project(DTYPE_F32)

    def _build_pca_search(self, components, explained: float, sample: int, chunk: int = 1 << 16):
        """PCA-compressed search: the candidates come from the dictionary's
        coordinates in its k leading principal directions (uncentred second
        moment of the zero-mean patterns, so the dot products themselves are
        approximated), in the search dtype; rescore_k > 0 re-scores them
        exactly against the full FP32 dictionary. The second moment is
        accumulated on the matrix cores (ebsd_cov_update: centred aggregate,
        plus n mean mean^T) over up to `sample` dictionary rows. Its leading
        eigenvectors come from FP64 subspace iteration (leading_eigh:
        products on the device, the small Rayleigh-Ritz solve on the host;
        basis_solver "subspace") or from the full FP64 solve on the host
        (fp64_eigh; "full", and the fallback when the leading 512 do not hold
        `explained`). components "auto": the smallest multiple of 64 (at
        least the search kernel's 4 K stages) whose leading eigenvalues hold
        `explained` of the trace; an int: that many."""
        from .pca_covmat import OnlineCovMatrix, fp64_eigh, leading_eigh

        full = self.dictionary_f32
        d = full.n_pix
        cov = OnlineCovMatrix(d, device=self.device)
        n_s = min(sample, full.rows)
        for s in range(0, n_s, chunk):
            cov(full.narrow(s, min(chunk, n_s - s)))
        n_obs = float(int(cov.obs.item()))
        mean = cov.mean.to(torch.float64)
        m2 = cov._full_aggregate().to(torch.float64) + n_obs * torch.outer(mean, mean)
        trace = float(torch.trace(m2))
        elem = kernels.TORCH_DTYPE[self.dtype].itemsize
        k_min = max(64, -(-256 // elem))  # >= 4 K stages of 64 bytes

        def rounded(k):
            return min(max(-(-k // 64) * 64, k_min), d)

        # one FP64 solve per dictionary
        if components == "auto":
            evals = evecs = None
            if self.basis_solver == "subspace":
                evals, evecs = leading_eigh(m2, rounded(512))
                if float(evals.clamp_min(0.0).sum()) < explained * trace:
                    evals = evecs = None  # the leading 512 fall short: full solve
            if evals is None:
                evals, evecs = fp64_eigh(m2)
            evals = evals.clamp_min(0.0)
            frac = torch.cumsum(evals.flip(0), 0) / trace
            k = rounded(int(torch.searchsorted(frac, torch.tensor([explained], dtype=frac.dtype,
                                                                   device=frac.device))) + 1)
        else:
            k = rounded(int(components))
            evals, evecs = leading_eigh(m2, k) if self.basis_solver == "subspace" else fp64_eigh(m2)
            evals = evals.clamp_min(0.0)
        self.search_components = k
        self.explained = float(evals[-k:].sum() / trace)
        basis = evecs[:, -k:]  # (d, k) float64
        if self.rotate_basis:
            # a fixed random rotation of the subspace: the dot products are
            # unchanged, the energy is spread evenly over the k coordinates,
            # so the low-precision rounding errors of the coordinates average
            # over k comparable terms instead of the few leading ones
            g = torch.Generator().manual_seed(0x5eed)
            rot, _ = torch.linalg.qr(torch.randn((k, k), generator=g, dtype=torch.float64))
            basis = basis @ rot.to(basis.device)
        basis = basis.to(torch.float32)
        bt = kernels.alloc_patterns(k, d, DTYPE_F32, self.device)
        bt.data.zero_()
        bt.data[:, :d] = basis.T
        self.basis_t = bt
        # FP8 search: the query coordinates are rounded to FP8 right after the
        # projection, so it runs on BF16 operands (half the bytes, the BF16
        # matrix cores); the dictionary coordinates above stay FP32-exact
        self.basis_t_bf16 = (kernels.prep_patterns(basis.T.contiguous(), DTYPE_BF16, zero_mean=False)
                             if self.dtype == DTYPE_FP8_E4M3 else None)
        coef = torch.empty((full.rows, kernels.padded_width(k, DTYPE_F32)), dtype=torch.float32, device=self.device)
        for s in range(0, full.rows, chunk):
            kernels.gemm_nt(full.narrow(s, min(chunk, full.rows - s)), bt, out=coef[s:s + chunk])
        self.dictionary = kernels.prep_patterns(coef[:, :k], self.dtype, zero_mean=False)
        del coef

    def prepare_pair(self, patterns: torch.Tensor, pix_idx: Optional[torch.Tensor] = None, zero_mean: bool = True,
                     unit_norm: bool = False):
        """(search rows, FP32 rows for rescoring or None) of (N, P) patterns
        (unit_norm: also divide each row by its norm, for cosine scores):
        with a PCA-compressed search the search rows are the patterns'
        coordinates in the basis, in the search dtype."""
        if self.basis_t is None:
            q = kernels.prep_patterns(patterns, self.dtype, zero_mean=zero_mean, unit_norm=unit_norm, pix_idx=pix_idx)
            qf = None
            if self.dictionary_f32 is self.dictionary:
                qf = q
            elif self.dictionary_f32 is not None:
                qf = kernels.prep_patterns(patterns, DTYPE_F32, zero_mean=zero_mean, unit_norm=unit_norm,
                                           pix_idx=pix_idx)
            return q, qf
        k = self.search_components
        if self.basis_t_bf16 is not None:
            qf, qb = kernels.prep_patterns_pair(patterns, DTYPE_BF16, zero_mean=zero_mean, unit_norm=unit_norm,
                                                pix_idx=pix_idx)
            qc = torch.empty((qf.rows, kernels.padded_width(k, DTYPE_F32)), dtype=torch.float32, device=self.device)
            kernels.gemm_nt(qb, self.basis_t_bf16, out=qc)
            del qb
        else:
            qf = kernels.prep_patterns(patterns, DTYPE_F32, zero_mean=zero_mean, unit_norm=unit_norm,
                                       pix_idx=pix_idx)
            qc = torch.empty((qf.rows, kernels.padded_width(k, DTYPE_F32)), dtype=torch.float32, device=self.device)
            kernels.gemm_nt(qf, self.basis_t, out=qc)
        return kernels.prep_patterns(qc[:, :k], self.dtype, zero_mean=False), qf

    def _project(self, dtype: int) -> kernels.PatternMatrix:
        return kernels.project(self.mp, self.rays, self.orientations, dtype, zero_mean=True, quads=self.quads,
                               unit_norm=self.unit_norm, renorm_dirs=self.renorm_dirs)

    @property
    def size(self) -> int:
        return self.orientations.shape[0]

    def prepare(self, patterns: torch.Tensor, pix_idx: Optional[torch.Tensor] = None,
                zero_mean: bool = True, dtype: Optional[int] = None) -> kernels.PatternMatrix:
        """(N, P) float patterns -> zero-mean padded rows of the search dtype
        (`dtype` given: full-width rows of that dtype). With a PCA-compressed
        search the search rows are the basis coordinates (prepare_pair gives
        both at once)."""
        if dtype is None and self.basis_t is not None:
            return self.prepare_pair(patterns, pix_idx, zero_mean)[0]
        return kernels.prep_patterns(patterns, self.dtype if dtype is None else dtype, zero_mean=zero_mean,
                                     pix_idx=pix_idx)



class PCADictionaryIndexer(_ResidentSearch):
    """PCA dictionary indexing engine (indexing.py:214-476).

    The covariance of the zero-mean learning-dictionary patterns is
    accumulated on the matrix cores (OnlineCovMatrix), its eigenvectors are
    solved in FP64, the chosen components (columns, largest eigenvalue last)
    form P (D x n_components), and the index dictionary is projected and
    stored as its PCA coefficients (N x n_components) in HBM. Queries are
    zero-meaned, projected onto P and searched like plain DI.

    coords are the un-normalised sample-frame pixel positions: the rotated
    directions are normalised inside the projection kernel, as
    interpolate(normalize_coords=True) does (indexing.py:323-335).
    component_slice "top" keeps the n largest components (the documented
    intent); "reference" reproduces indexing.py:355's slice [-n : -n + 300].
    """

    def __init__(self, master_pattern: torch.Tensor, coords: torch.Tensor, learn_orientations: torch.Tensor,
                 n_components: int, index_orientations: Optional[torch.Tensor] = None, precision: str = "fp32",
                 rescore_k: int = 0, component_slice: str = "top", chunk: int = 1 << 16):
        from .pca_covmat import OnlineCovMatrix

        if precision not in kernels.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(kernels.PRECISIONS)}")
        self.device = learn_orientations.device
        self.mp = master_pattern.to(self.device, torch.float32).contiguous()
        self.coords = coords.to(self.device, torch.float32).contiguous()
        self.quads = kernels.mp_quad_table(self.mp)
        self.precision = precision
        self.dtype = kernels.PRECISIONS[precision]
        self.rescore_k = int(rescore_k)
        d = self.coords.shape[0]
        self.n_pix = d
        learn = learn_orientations.to(self.device, torch.float32).contiguous()
        self.covariance = OnlineCovMatrix(d, device=self.device)
        # stage times (milliseconds) of the build, for measurement: the
        # learning projection, the covariance updates, the FP64 eigensolve,
        # and the index dictionary's projection onto the components
        self.timings = {"learn_projection_ms": 0.0, "covariance_ms": 0.0}
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        for s in range(0, learn.shape[0], chunk):
            ev[0].record()
            x = self._project(learn[s:s + chunk])
            ev[1].record()
            self.covariance(x)
            ev[2].record()
            ev[2].synchronize()
            self.timings["learn_projection_ms"] += ev[0].elapsed_time(ev[1])
            self.timings["covariance_ms"] += ev[1].elapsed_time(ev[2])
        t0 = time.perf_counter()
        self.eigenvalues, evecs = self.covariance.get_eigensystem(torch.float64)
        self.timings["eigensolve_ms"] = (time.perf_counter() - t0) * 1000.0
        if component_slice == "top":
            comps = evecs[:, -n_components:]
        elif component_slice == "reference":
            comps = evecs[:, -n_components:(-n_components + 300)]
        else:
            raise ValueError("component_slice must be 'top' or 'reference'")
        if comps.shape[1] == 0:
            raise ValueError("the component slice is empty")
        self.components = comps  # (D, n) float64
        n = comps.shape[1]
        self.n_components = n
        pt = kernels.alloc_patterns(kernels.padded_width(n, DTYPE_F32), d, DTYPE_F32, self.device)
        pt.data.zero_()
        pt.data[:n, :d] = comps.T.to(torch.float32)
        self.components_t = pt
        index = learn if index_orientations is None else index_orientations.to(self.device, torch.float32)
        self.orientations = index.contiguous()
        n_pad = kernels.padded_width(n, DTYPE_F32)
        f32 = torch.empty((index.shape[0], n_pad), dtype=torch.float32, device=self.device)
        self.timings["index_projection_ms"] = 0.0
        self.timings["index_coefficients_ms"] = 0.0
        for s in range(0, index.shape[0], chunk):
            ev[0].record()
            x = self._project(index[s:s + chunk])
            ev[1].record()
            kernels.gemm_nt(x, pt, out=f32[s:s + chunk])
            ev[2].record()
            ev[2].synchronize()
            self.timings["index_projection_ms"] += ev[0].elapsed_time(ev[1])
            self.timings["index_coefficients_ms"] += ev[1].elapsed_time(ev[2])
        dict_f32 = kernels.PatternMatrix(f32, None, DTYPE_F32, n)
        if self.dtype == DTYPE_F32:
            # exact FP32 re-ranking of the matrix-core search's best
            # candidates on the coefficients (see DictionaryIndexer)
            self.dictionary, self.dictionary_f32 = dict_f32, dict_f32
            self.rescore_k = EBSD_MAX_TOPK
        else:
            self.dictionary = kernels.prep_patterns(f32[:, :n], self.dtype, zero_mean=False)
            self.dictionary_f32 = dict_f32 if self.rescore_k > 0 else None

    def _project(self, quats: torch.Tensor) -> kernels.PatternMatrix:
        return kernels.project(self.mp, self.coords, quats, DTYPE_F32, zero_mean=True, renorm_dirs=True,
                               quads=self.quads)

    @property
    def size(self) -> int:
        return self.orientations.shape[0]

    def prepare_pair(self, patterns: torch.Tensor, pix_idx: Optional[torch.Tensor] = None, zero_mean: bool = True):
        """(search rows, FP32 coefficient rows for rescoring or None)."""
        q = self.prepare(patterns, pix_idx, zero_mean)
        if self.dictionary_f32 is self.dictionary:
            return q, q
        if self.dictionary_f32 is not None:
            return q, self.prepare(patterns, pix_idx, zero_mean, DTYPE_F32)
        return q, None

    def prepare(self, patterns: torch.Tensor, pix_idx: Optional[torch.Tensor] = None,
                zero_mean: bool = True, dtype: Optional[int] = None) -> kernels.PatternMatrix:
        """(N, P) patterns -> PCA coefficients (N, n_components) in the search
        dtype (or `dtype`)."""
        q = kernels.prep_patterns(patterns, DTYPE_F32, zero_mean=zero_mean, pix_idx=pix_idx)
        n_pad = self.components_t.rows
        out = torch.empty((q.rows, n_pad), dtype=torch.float32, device=self.device)
        kernels.gemm_nt(q, self.components_t, out=out)
        pm = kernels.PatternMatrix(out, None, DTYPE_F32, self.n_components)
        dt = self.dtype if dtype is None else dtype
        if dt == DTYPE_F32:
            return pm
        return kernels.prep_patterns(out[:, :self.n_components], dt, zero_mean=False)


def dictionary_index_orientations(
    master_patterns: Union[MasterPattern, List[MasterPattern]],
    geometry: EBSDGeometry,
    experiment_patterns: ExperimentPatterns,
    signal_mask: Optional[torch.Tensor] = None,
    subtract_exp_pat_mean: bool = True,
    experiment_chunk_size: int = 256,
    dictionary_resolution_degrees: float = 0.5,
    dictionary_chunk_size: int = 256,
    virtual_binning: int = 1,
    top_k_matches: int = 1,
    distance_metric: str = "angular",
    match_dtype: torch.dtype = torch.float16,
    quantized_if_x86: bool = True,
    average_pattern_center: bool = True,
    *,
    dictionary_orientations: Optional[Sequence[torch.Tensor]] = None,
    precision: Optional[str] = None,
    rescore_k: int = 0,
    dict_unit_norm: bool = False,
    generator: Optional[torch.Generator] = None,
    search_components: Union[int, str] = 0,
) -> None:
    """Index experimental patterns against a projected dictionary per phase
    (indexing.py:14-211) and store the results in `experiment_patterns`:
    raw_indexing_results[phase] = (quaternions (N, k, 4), metrics (N, k)),
    raw_indexing_indices[phase] = dictionary indices (N, k) int64, then
    phases / orientations / best_metrics via combine_indexing_results.

    Extensions (keyword-only, defaults keep the reference semantics):
    dictionary_orientations  per-phase (N_dict, 4) dictionaries to use
                             instead of sampling one
    precision                "fp32" | "bf16" | "fp16" | "fp8"; default
                             follows match_dtype
    rescore_k                re-score this many low-precision candidates in
                             FP32 (precision form)
    dict_unit_norm           rank by cosine instead of the raw dot product
    generator                seeds the dictionary permutation
    search_components        with rescore_k > 0: search on the dictionary's
                             leading principal coordinates ("auto" or an
                             int; 0 = full patterns), then rescore exactly
    """
    if isinstance(master_patterns, MasterPattern):
        master_patterns = [master_patterns]
    if distance_metric != "angular":
        raise NotImplementedError(f"'{distance_metric}' not supported.")
    if not average_pattern_center:
        raise NotImplementedError("per-scan-point pattern centres are not supported")
    if precision is None:
        if match_dtype not in kernels.MATCH_DTYPES:
            raise NotImplementedError(f"match_dtype {match_dtype} is not supported")
        precision = {torch.float32: "fp32", torch.bfloat16: "bf16", torch.float16: "fp16"}[match_dtype]
    n_exp = experiment_patterns.n_patterns
    for i, mp in enumerate(master_patterns):
        device = _gpu_device(mp.master_pattern)
        if dictionary_orientations is not None:
            ori = dictionary_orientations[i].to(device, torch.float32)
        else:
            ori = sample_ori_fz_laue_angle(mp.laue_group, dictionary_resolution_degrees, device, permute=True,
                                           generator=generator)
        rays = geometry.get_rays((virtual_binning, virtual_binning))
        pix_idx = None
        if signal_mask is not None:
            flat = signal_mask.flatten().to(rays.device)
            rays = rays[flat]
            pix_idx = torch.nonzero(flat.to(device), as_tuple=False).flatten().to(torch.int32)
        engine = DictionaryIndexer(mp.pattern_for_binning(virtual_binning), rays.to(device), ori,
                                   precision=precision, rescore_k=rescore_k, dict_unit_norm=dict_unit_norm,
                                   search_components=search_components)
        exp = experiment_patterns.get_patterns(torch.arange(n_exp, device=experiment_patterns.patterns.device),
                                               binning=virtual_binning)
        exp = exp.reshape(n_exp, -1).to(device, torch.float32)
        # the whole-set query path of the reference honours
        # subtract_exp_pat_mean; its chunked path always subtracts
        zero_mean = subtract_exp_pat_mean if experiment_chunk_size > n_exp else True
        topk = top_k_matches
        vals, idxs = [], []
        # bound the prepared query buffer to ~4 GB per step
        step = max(1, min(n_exp, (4 << 30) // (4 * max(1, exp.shape[1]))))
        for s in range(0, n_exp, step):
            q, qf = engine.prepare_pair(exp[s:s + step], pix_idx, zero_mean)
            v, ix = engine.index(q, topk, qf)
            vals.append(v)
            idxs.append(ix)
        val = torch.cat(vals, 0)
        idx = torch.cat(idxs, 0).to(torch.int64)
        if match_dtype in (torch.float16, torch.bfloat16) and precision in ("fp16", "bf16") and rescore_k == 0:
            val = val.to(match_dtype)
        experiment_patterns.set_raw_indexing_results(ori[idx.clamp_min(0)], val, phase_id=i, indices=idx)
    experiment_patterns.combine_indexing_results(higher_is_better=True)


def pca_dictionary_index_orientations(
    master_patterns: Union[MasterPattern, List[MasterPattern]],
    geometry: EBSDGeometry,
    experiment_patterns: ExperimentPatterns,
    n_pca_components: int,
    signal_mask: Optional[torch.Tensor] = None,
    subtract_exp_pat_mean: bool = True,
    experiment_chunk_size: int = 256,
    dictionary_resolution_learn_deg: float = 0.5,
    dictionary_resolution_index_deg: float = 0.5,
    dictionary_chunk_size: int = 256,
    virtual_binning: int = 1,
    top_k_matches: int = 1,
    distance_metric: str = "angular",
    match_dtype: torch.dtype = torch.float16,
    quantized_via_ao: bool = True,
    average_pattern_center: bool = True,
    *,
    learn_orientations: Optional[Sequence[torch.Tensor]] = None,
    index_orientations: Optional[Sequence[torch.Tensor]] = None,
    precision: Optional[str] = None,
    rescore_k: int = 0,
    component_slice: str = "top",
    generator: Optional[torch.Generator] = None,
) -> None:
    """PCA dictionary indexing (indexing.py:214-476): learn the covariance
    of a learning dictionary, keep n_pca_components principal components,
    index in PCA space; results are stored in `experiment_patterns` as by
    dictionary_index_orientations.

    Extensions (keyword-only): learn_orientations / index_orientations
    (per-phase dictionaries instead of sampling), precision ("fp32" |
    "bf16" | "fp16" | "fp8", default from match_dtype), rescore_k (FP32 re-scoring in
    PCA space), component_slice ("top" = the n largest components, the
    documented intent; "reference" = the literal slice of indexing.py:355).
    """
    if isinstance(master_patterns, MasterPattern):
        master_patterns = [master_patterns]
    if distance_metric != "angular":
        raise NotImplementedError(f"'{distance_metric}' not supported.")
    if not average_pattern_center:
        raise NotImplementedError("per-scan-point pattern centres are not supported")
    if precision is None:
        if match_dtype not in kernels.MATCH_DTYPES:
            raise NotImplementedError(f"match_dtype {match_dtype} is not supported")
        precision = {torch.float32: "fp32", torch.bfloat16: "bf16", torch.float16: "fp16"}[match_dtype]
    n_exp = experiment_patterns.n_patterns
    for i, mp in enumerate(master_patterns):
        device = _gpu_device(mp.master_pattern)
        if learn_orientations is not None:
            learn = learn_orientations[i].to(device, torch.float32)
        else:
            learn = sample_ori_fz_laue_angle(mp.laue_group, dictionary_resolution_learn_deg, device, permute=True,
                                             generator=generator)
        if index_orientations is not None:
            index = index_orientations[i].to(device, torch.float32)
        elif dictionary_resolution_learn_deg != dictionary_resolution_index_deg:
            index = sample_ori_fz_laue_angle(mp.laue_group, dictionary_resolution_index_deg, device, permute=True,
                                             generator=generator)
        else:
            index = learn
        coords = geometry.get_coords_sample_frame((virtual_binning, virtual_binning))
        pix_idx = None
        if signal_mask is not None:
            flat = signal_mask.flatten().to(coords.device)
            coords = coords[flat]
            pix_idx = torch.nonzero(flat.to(device), as_tuple=False).flatten().to(torch.int32)
        engine = PCADictionaryIndexer(mp.pattern_for_binning(virtual_binning), coords.to(device), learn,
                                      n_pca_components, index_orientations=index, precision=precision,
                                      rescore_k=rescore_k, component_slice=component_slice)
        exp = experiment_patterns.get_patterns(torch.arange(n_exp, device=experiment_patterns.patterns.device),
                                               binning=virtual_binning)
        exp = exp.reshape(n_exp, -1).to(device, torch.float32)
        zero_mean = subtract_exp_pat_mean if experiment_chunk_size > n_exp else True
        vals, idxs = [], []
        step = max(1, min(n_exp, (4 << 30) // (4 * max(1, exp.shape[1]))))
        for s in range(0, n_exp, step):
            q, qf = engine.prepare_pair(exp[s:s + step], pix_idx, zero_mean)
            v, ix = engine.index(q, top_k_matches, qf)
            vals.append(v)
            idxs.append(ix)
        val = torch.cat(vals, 0)
        idx = torch.cat(idxs, 0).to(torch.int64)
        experiment_patterns.set_raw_indexing_results(engine.orientations[idx.clamp_min(0)], val, phase_id=i,
                                                     indices=idx)
    experiment_patterns.combine_indexing_results(higher_is_better=True)
