"""Build libgnss_gpu_amd.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m gnss_gpu_amd.build  [--arch gfx950] [--jobs N]
"""

from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
INCLUDE = os.path.join(REPO_DIR, "include")
LIB_NAME = "libgnss_gpu_amd.so"
SOURCES = ["gg_common.hip", "gg_pf.hip", "gg_wls.hip", "gg_learn.hip", "gg_bvh.hip",
           "gg_smooth.hip"]


def hipcc_path() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build gnss_gpu_amd)")


def compile_flags(arch: str) -> list[str]:
    return [
        f"--offload-arch={arch}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-I" + INCLUDE,
        "-I" + CSRC,
        # binary32 packed math is kept off (compiler-formed v_pk_* only)
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        # matrix-core results straight into VGPRs (no accvgpr copies)
        "-mllvm", "-amdgpu-mfma-vgpr-form",
        "-Wno-unused-command-line-argument",
    ]


def build(arch: str = "gfx950", jobs: int = 4, verbose: bool = False) -> str:
    hipcc = hipcc_path()
    obj_dir = os.path.join(PKG_DIR, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    flags = compile_flags(arch)

    def one(src: str) -> str:
        obj = os.path.join(obj_dir, src.replace(".hip", ".o"))
        cmd = [hipcc, *flags, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
        return obj

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(one, SOURCES))
    out = os.path.join(PKG_DIR, LIB_NAME)
    tmp = out + ".tmp"
    cmd = [hipcc, f"--offload-arch={arch}", "-shared", "-fPIC", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--arch", default=os.environ.get("PYTORCH_ROCM_ARCH", "gfx950"))
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("-v", "--verbose", action="store_true")
    args = ap.parse_args(argv)
    path = build(args.arch, args.jobs, args.verbose)
    print(path)
    return 0


if __name__ == "__main__":
    sys.exit(main())
