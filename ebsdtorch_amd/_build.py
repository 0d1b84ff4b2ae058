"""Build recipe of libebsd_amd.so (gfx950) from ebsdtorch_amd/csrc.

Each HIP source is compiled to an object with hipcc and the objects are
linked into one shared library next to this file, so the library ships with
the source tree. Packed FP32 instructions are disabled for all device code.
"""

from __future__ import annotations

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_NAME = "libebsd_amd.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)
ARCH = "gfx950"

SOURCES = {
    "abi.hip": [],
    "project.hip": [],
    "simtopk.hip": [],
    "pca.hip": [],
    "noise.hip": [],
    "nn.hip": [],
    "local.hip": [],
    "nlpar.hip": [],
    "preprocess.hip": [],
    "fitting.hip": [],
}

COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc_path() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def _compile(args):
    cmd, src = args
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass reports the device-only feature flag; not an error.
    noise = "'-packed-fp32-ops' is not a recognized feature"
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    return err


def build(verbose: bool = False, jobs: int = 4) -> str:
    """Compile every HIP source for gfx950 and link LIB_PATH. Returns it."""
    hipcc = hipcc_path()
    obj_dir = os.path.join(PKG_DIR, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    jobs_list = []
    objs = []
    for src, extra in SOURCES.items():
        obj = os.path.join(obj_dir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [hipcc, *COMMON_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
        jobs_list.append((cmd, src))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for (cmd, src), warn in zip(jobs_list, ex.map(_compile, jobs_list)):
            if verbose:
                print(" ".join(cmd))
                if warn:
                    print(warn)
    tmp = LIB_PATH + ".tmp"
    link = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(link, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(" ".join(link))
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
