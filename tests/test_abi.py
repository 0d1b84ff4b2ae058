"""The C ABI library loads, exports every symbol the header declares, and
rejects bad arguments with status codes before touching the GPU."""

import ctypes
import os
import re

import pytest

from ebsdtorch_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ebsd_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ebsd_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_lib.exported_symbols())


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert lib.ebsd_abi_version() == _lib.EBSD_ABI_VERSION


def test_pad_multiples():
    lib = _lib.load()
    assert lib.ebsd_pad_multiple(_lib.DTYPE_F32) == 32
    assert lib.ebsd_pad_multiple(_lib.DTYPE_BF16) == 64
    assert lib.ebsd_pad_multiple(_lib.DTYPE_F16) == 64
    assert lib.ebsd_pad_multiple(_lib.DTYPE_FP8_E4M3) == 128
    assert lib.ebsd_pad_multiple(99) == _lib.EBSD_ERR_DTYPE


def test_argument_errors_return_codes_without_launching():
    lib = _lib.load()
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: every call below fails validation first
    assert lib.ebsd_project(None, 5, 10, fake, 4, fake, 1, fake, 0, 32, None, 1, None) == _lib.EBSD_ERR_NULL
    assert lib.ebsd_project(fake, 5, 0, fake, 4, fake, 1, fake, 0, 32, None, 1, None) == _lib.EBSD_ERR_SHAPE
    assert lib.ebsd_project(fake, 5, 10, fake, 40, fake, 1, fake, 0, 32, None, 1, None) == _lib.EBSD_ERR_SHAPE
    assert lib.ebsd_project(fake, 5, 10, fake, 4, fake, 1, fake, 0, 33, None, 1, None) == _lib.EBSD_ERR_SHAPE
    assert lib.ebsd_project(fake, 5, 10, fake, 4, fake, 1, fake, 7, 32, None, 1, None) == _lib.EBSD_ERR_DTYPE
    assert lib.ebsd_project(fake, 5, 10, fake, 4, fake, 1, fake, 3, 128, None, 1, None) == _lib.EBSD_ERR_NULL
    assert lib.ebsd_project(fake, 5, 10, fake, 4, ctypes.c_void_p(0x1004), 1, fake, 0, 32, None, 1,
                            None) == _lib.EBSD_ERR_ALIGN
    assert lib.ebsd_prep_patterns(None, 1, 4, None, 4, fake, 0, 32, None, 1, None) == _lib.EBSD_ERR_NULL
    assert lib.ebsd_prep_patterns(fake, 1, 4, None, 8, fake, 0, 32, None, 1, None) == _lib.EBSD_ERR_SHAPE
    assert lib.ebsd_prep_patterns_pair(None, 1, 4, None, 4, fake, 32, fake, 1, 64, 1, None) == _lib.EBSD_ERR_NULL
    assert lib.ebsd_prep_patterns_pair(fake, 1, 4, None, 4, fake, 32, fake, 3, 128, 1, None) == _lib.EBSD_ERR_DTYPE
    assert lib.ebsd_prep_patterns_pair(fake, 1, 4, None, 4, fake, 32, fake, 1, 32, 1, None) == _lib.EBSD_ERR_SHAPE
    assert lib.ebsd_prep_patterns_pair(fake, 1, 4, None, 8, fake, 32, fake, 1, 64, 1, None) == _lib.EBSD_ERR_SHAPE
    assert lib.ebsd_sample_dirs(fake, 5, 10, None, 3, fake, 0, None) == _lib.EBSD_ERR_NULL
    nb = ctypes.c_size_t(0)
    assert lib.ebsd_sim_topk_workspace(10, 100, 17, 0, ctypes.byref(nb)) == _lib.EBSD_ERR_TOPK
    assert lib.ebsd_sim_topk_workspace(10, 100, 0, 0, ctypes.byref(nb)) == _lib.EBSD_ERR_TOPK
    assert lib.ebsd_sim_topk_workspace(10, 100, 4, 0, ctypes.byref(nb)) == _lib.EBSD_OK and nb.value > 0
    args = [fake, None, 10, fake, None, 100, 32, 0, 4, fake, fake, 0, 0, fake, 0, None]
    assert lib.ebsd_sim_topk(*args) == _lib.EBSD_ERR_WORKSPACE
    bad = list(args)
    bad[6] = 33
    assert lib.ebsd_sim_topk(*bad) == _lib.EBSD_ERR_SHAPE
    bad = list(args)
    bad[0] = None
    assert lib.ebsd_sim_topk(*bad) == _lib.EBSD_ERR_NULL
    bad = list(args)
    bad[11] = 2 ** 31 - 50
    assert lib.ebsd_sim_topk(*bad) == _lib.EBSD_ERR_SHAPE
    assert lib.ebsd_rescore(fake, 4, fake, 9, 32, fake, 17, 1, fake, fake, None) == _lib.EBSD_ERR_TOPK
    assert lib.ebsd_rescore(fake, 4, fake, 9, 32, fake, 4, 5, fake, fake, None) == _lib.EBSD_ERR_TOPK
    assert lib.ebsd_rescore(fake, 4, fake, 0, 32, fake, 4, 2, fake, fake, None) == _lib.EBSD_ERR_SHAPE
    assert b"workspace" in lib.ebsd_status_string(_lib.EBSD_ERR_WORKSPACE)


def test_status_codes_map_to_reference_exception_types():
    with pytest.raises(ValueError):
        _lib.check(_lib.EBSD_ERR_SHAPE, "x")
    with pytest.raises(NotImplementedError):
        _lib.check(_lib.EBSD_ERR_DTYPE, "x")
    with pytest.raises(_lib.EbsdLibraryError):
        _lib.check(700, "x")
