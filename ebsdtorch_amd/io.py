"""EMsoft master-pattern files (io/read_master_pattern.py:7-53).

read_master_pattern(file_path) -> MasterPattern, as the reference: the
Monte-Carlo energy histogram EMData/MCOpenCL/accum_e summed over its first
two axes gives the energy pdf, which weights the per-energy square-Lambert
master patterns EMData/EBSDmaster/mLPNH and mLPSH; the two hemispheres are
placed side by side (north left) and CrystalData/SpaceGroupNumber selects the
Laue group.

The reference reads the file with h5py, which this image does not have.
Here the file is read through the HDF5 C library itself (libhdf5, bound with
ctypes): each dataset is read into float64 (exact for EMsoft's integer
counts and float32 patterns) and cast to float32, the cast the reference
applies (astype(np.float32)), so the arithmetic that follows sees the same
values. The library is looked up in EBSD_HDF5_LIB, then by name, then in
the usual install locations; without it the reader raises OSError.
"""

from __future__ import annotations

import ctypes
import ctypes.util
import os
from typing import Dict, Optional

import numpy as np
import torch

from .master_pattern import MasterPattern

_hid = ctypes.c_int64  # hid_t of HDF5 >= 1.10
_hsize = ctypes.c_uint64
_H5F_ACC_RDONLY = 0
_H5P_DEFAULT = 0
_H5S_ALL = 0

_CANDIDATES = ("/opt/conda/lib/libhdf5.so", "/opt/conda/lib/libhdf5.so.103", "/usr/lib/x86_64-linux-gnu/libhdf5.so",
               "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so", "/usr/lib64/libhdf5.so", "/usr/local/lib/libhdf5.so")
_lib = None


def _hdf5():
    """Load (once) the HDF5 C library and declare the calls used here."""
    global _lib
    if _lib is not None:
        return _lib
    names = [os.environ.get("EBSD_HDF5_LIB"), ctypes.util.find_library("hdf5"), *_CANDIDATES]
    errors = []
    for name in names:
        if not name:
            continue
        try:
            lib = ctypes.CDLL(name)
        except OSError as e:
            errors.append(f"{name}: {e}")
            continue
        if not hasattr(lib, "H5Dopen2"):
            errors.append(f"{name}: no H5Dopen2 (HDF5 < 1.8?)")
            continue
        break
    else:
        raise OSError("the HDF5 C library (libhdf5) was not found; set EBSD_HDF5_LIB to its path. Tried: "
                      + "; ".join(errors or ["nothing on the search path"]))
    lib.H5open.restype = ctypes.c_int
    lib.H5Eset_auto2.argtypes = [_hid, ctypes.c_void_p, ctypes.c_void_p]
    lib.H5Fopen.restype = _hid
    lib.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, _hid]
    lib.H5Fclose.argtypes = [_hid]
    lib.H5Lexists.restype = ctypes.c_int
    lib.H5Lexists.argtypes = [_hid, ctypes.c_char_p, _hid]
    lib.H5Dopen2.restype = _hid
    lib.H5Dopen2.argtypes = [_hid, ctypes.c_char_p, _hid]
    lib.H5Dclose.argtypes = [_hid]
    lib.H5Dget_space.restype = _hid
    lib.H5Dget_space.argtypes = [_hid]
    lib.H5Sclose.argtypes = [_hid]
    lib.H5Sget_simple_extent_ndims.restype = ctypes.c_int
    lib.H5Sget_simple_extent_ndims.argtypes = [_hid]
    lib.H5Sget_simple_extent_dims.restype = ctypes.c_int
    lib.H5Sget_simple_extent_dims.argtypes = [_hid, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]
    lib.H5Dread.restype = ctypes.c_int
    lib.H5Dread.argtypes = [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]
    if lib.H5open() < 0:
        raise OSError("H5open failed")
    lib.H5Eset_auto2(0, None, None)  # errors come back as return codes, not printed
    _lib = lib
    return lib


class _File:
    def __init__(self, path: str):
        self.lib = _hdf5()
        self.path = path
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.fid = self.lib.H5Fopen(path.encode(), _H5F_ACC_RDONLY, _H5P_DEFAULT)
        if self.fid < 0:
            raise OSError(f"{path}: not an HDF5 file, or it cannot be opened")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.lib.H5Fclose(self.fid)

    def _exists(self, name: str) -> bool:
        # every component of the path must exist before the last is tested
        parts = name.strip("/").split("/")
        for i in range(1, len(parts) + 1):
            if self.lib.H5Lexists(self.fid, "/".join(parts[:i]).encode(), _H5P_DEFAULT) <= 0:
                return False
        return True

    def read(self, name: str) -> np.ndarray:
        """The whole dataset `name`, converted by HDF5 to float64."""
        lib = self.lib
        if not self._exists(name):
            raise KeyError(f"{self.path}: no dataset {name!r}")
        did = lib.H5Dopen2(self.fid, name.encode(), _H5P_DEFAULT)
        if did < 0:
            raise KeyError(f"{self.path}: {name!r} cannot be opened as a dataset")
        try:
            sid = lib.H5Dget_space(did)
            try:
                nd = lib.H5Sget_simple_extent_ndims(sid)
                if nd < 0:
                    raise OSError(f"{self.path}: {name!r} has no simple dataspace")
                dims = (_hsize * max(nd, 1))()
                lib.H5Sget_simple_extent_dims(sid, dims, None)
                shape = tuple(int(dims[i]) for i in range(nd))
            finally:
                lib.H5Sclose(sid)
            out = np.empty(shape, dtype=np.float64)
            mem_type = _hid.in_dll(lib, "H5T_NATIVE_DOUBLE_g").value
            if lib.H5Dread(did, mem_type, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, out.ctypes.data) < 0:
                raise OSError(f"{self.path}: reading {name!r} as numbers failed")
            return out
        finally:
            lib.H5Dclose(did)


def read_emsoft_arrays(file_path: str) -> Dict[str, np.ndarray]:
    """The datasets read_master_pattern uses, as float32 arrays (space group
    as a Python int): accum_e, mLPNH, mLPSH, space_group."""
    with _File(str(file_path)) as f:
        return {
            "accum_e": f.read("EMData/MCOpenCL/accum_e").astype(np.float32),
            "mLPNH": f.read("EMData/EBSDmaster/mLPNH").astype(np.float32),
            "mLPSH": f.read("EMData/EBSDmaster/mLPSH").astype(np.float32),
            "space_group": int(f.read("CrystalData/SpaceGroupNumber").reshape(-1)[0]),
        }


def energy_weighted_master_pattern(accum_e: np.ndarray, mlpnh: np.ndarray, mlpsh: np.ndarray) -> np.ndarray:
    """Side-by-side (H, 2W) float32 master pattern: the per-energy square
    Lambert patterns (1, n_energies, H, W) weighted by the energy pdf of the
    Monte-Carlo histogram (its sums over the first two axes, normalised),
    the arithmetic of read_master_pattern.py:36-47 in float32."""
    north = np.asarray(mlpnh, dtype=np.float32).squeeze()
    south = np.asarray(mlpsh, dtype=np.float32).squeeze()
    probabilities = np.asarray(accum_e, dtype=np.float32).sum(axis=(0, 1))
    pdf = probabilities / probabilities.sum()
    nh = (pdf[:, None, None] * north).sum(axis=0)
    sh = (pdf[:, None, None] * south).sum(axis=0)
    return np.concatenate([nh, sh], axis=-1)


def read_master_pattern(file_path, device: Optional[torch.device] = None) -> MasterPattern:
    """Read an EMsoft EBSD master-pattern file into a MasterPattern
    (io/read_master_pattern.py:7-53); `device` places its buffer (default:
    the CPU, as the reference)."""
    arrays = read_emsoft_arrays(file_path)
    mp = energy_weighted_master_pattern(arrays["accum_e"], arrays["mLPNH"], arrays["mLPSH"])
    t = torch.tensor(mp)
    if device is not None:
        t = t.to(device)
    return MasterPattern(t, space_group=arrays["space_group"])
