"""Builds liblsm_hip.so (the HIP kernels and the C ABI) for gfx950 with hipcc.

    python -m lsm_amd.build        (from ns-lsm-single-phase_amd/)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB = os.path.join(PKG, "liblsm_hip.so")
SOURCES = ["lsm_capi.hip", "lsm_advect.hip", "lsm_reinit.hip", "lsm_volume.hip", "lsm_host.hip",
           "lsm_adjoint.hip", "lsm_reinit_adjoint.hip", "lsm_batch.hip", "lsm_consumers.hip",
           "lsm_upwind.hip", "lsm_geometric.hip", "lsm_multi.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LSM_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17",
         # reference operation order: no implicit multiply-add contraction
         "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Werror=infinite-recursion",
         "-Werror=uninitialized",
         f"-I{INCLUDE}", f"-I{CSRC}"]


def build(verbose: bool = False) -> str:
    obj_dir = os.path.join(ROOT, "_build")
    os.makedirs(obj_dir, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(obj_dir, src.replace(".hip", ".o"))
        cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    shutil.move(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
