"""The allocation-free similarity entry point (ebsd_sim_topk_noalloc),
called through the C ABI: its results equal ebsd_sim_topk's (same kernels,
static scratch instead of a caller workspace) bit for bit, on two calls in a
row with the second one smaller (the buffer keeps the first call's
contents), for the FP32 and the FP8 kernels; sizes beyond the static buffer
are refused before any device work."""

import ctypes

import pytest
import torch

from ebsdtorch_amd import _lib, kernels
from ebsdtorch_amd._lib import DTYPE_F32, DTYPE_FP8_E4M3

DEV = torch.device("cuda")


def test_oversize_is_refused_without_device_work():
    lib = _lib.load()
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: the size check fails first
    # FP32, top-16: 8 partial lists x 16 x 8 bytes per query -> 131,072
    # queries need 128 MiB of lists, twice the static buffer
    st = lib.ebsd_sim_topk_noalloc(fake, None, 131072, fake, None, 4096, 3616, DTYPE_F32, 16, fake, fake, 0, 0,
                                   None)
    assert st == _lib.EBSD_ERR_SCRATCH
    assert b"static scratch" in lib.ebsd_status_string(st)
    assert lib.ebsd_sim_topk_noalloc(None, None, 4, fake, None, 4, 32, DTYPE_F32, 1, fake, fake, 0, 0,
                                     None) == _lib.EBSD_ERR_NULL
    assert lib.ebsd_sim_topk_noalloc(fake, None, 4, fake, None, 4, 32, DTYPE_F32, 17, fake, fake, 0, 0,
                                     None) == _lib.EBSD_ERR_TOPK


def _run_both(q, d, k, merge_state=None):
    m = q.rows
    outs = []
    for fn in (kernels.sim_topk, kernels.sim_topk_noalloc):
        if merge_state is None:
            v = torch.empty((m, k), dtype=torch.float32, device=DEV)
            i = torch.empty((m, k), dtype=torch.int32, device=DEV)
            fn(q, d, k, v, i, index_base=0, merge=False)
        else:
            v, i = merge_state[0].clone(), merge_state[1].clone()
            fn(q, d, k, v, i, index_base=1000, merge=True)
        outs.append((v, i))
    torch.cuda.synchronize()
    return outs


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [DTYPE_F32, DTYPE_FP8_E4M3])
def test_noalloc_equals_workspace_entry_point(dtype):
    print(f"noalloc dtype={dtype}", flush=True)
    g = torch.Generator().manual_seed(40 + dtype)
    D = 1000
    dict_raw = torch.randn(5000, D, generator=g).to(DEV)
    d = kernels.prep_patterns(dict_raw, dtype, zero_mean=True)
    for m, k in ((3000, 16), (700, 5)):  # second call smaller
        q = kernels.prep_patterns(torch.randn(m, D, generator=g).to(DEV), dtype, zero_mean=True)
        (v0, i0), (v1, i1) = _run_both(q, d, k)
        assert torch.equal(i0, i1) and torch.equal(v0, v1), (m, k)
        assert torch.all(v0[:, 1:] <= v0[:, :-1])
        # merging a second dictionary chunk into the running state
        d2 = kernels.prep_patterns(torch.randn(600, D, generator=g).to(DEV), dtype, zero_mean=True)
        (w0, j0), (w1, j1) = _run_both(q, d2, k, merge_state=(v0, i0))
        assert torch.equal(j0, j1) and torch.equal(w0, w1), (m, k, "merge")
    # an exact FP32 check of the last call against float64 scores
    if dtype == DTYPE_F32:
        ref = q.to_float().double() @ d.to_float().double().T
        best = ref.max(1).values
        assert torch.allclose(v1[:, 0].double(), best, rtol=1e-5, atol=1e-4)
