"""Build the HIP library ``libfl_correction_amd.so`` in-tree for gfx950.

    python -m fl_correction_amd.build

Compiles every ``csrc/*.hip`` with hipcc (one object per source, in
parallel) and links them into ``fl_correction_amd/lib/``.  The library exports
only the C ABI of ``include/fl_correction_amd.h``.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_DIR = os.path.join(PKG, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libfl_correction_amd.so")
ARCH = os.environ.get("FL_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    # Packed FP32 instructions are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]

# Per-source additions.  bwd_sm.hip dispatches every hit to the registers of
# its cell through a tree of wave-uniform branches; the control-flow
# structurizer would otherwise rewrite that tree into flow blocks whose merges
# copy every accumulator at every hit (about three register moves per
# multiply-add).  Uniform regions need no structurizing: the branches are
# scalar.
EXTRA_FLAGS = {
    "bwd_sm.hip": ["-mllvm", "-structurizecfg-skip-uniform-regions=true"],
}


def sources():
    return sorted(glob.glob(os.path.join(PKG, "csrc", "*.hip")))


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build(verbose: bool = False, jobs: int = 8) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    with tempfile.TemporaryDirectory(prefix="fl_build_") as obj_dir:
        def compile_one(src):
            obj = os.path.join(obj_dir, os.path.basename(src) + ".o")
            cmd = [hipcc(), *HIPCC_FLAGS, *EXTRA_FLAGS.get(os.path.basename(src), []),
                   "-I", os.path.join(ROOT, "include"), "-c", src, "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True)
            return obj

        with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(sources())))) as pool:
            objs = list(pool.map(compile_one, sources()))
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-fPIC", "-shared", *objs, "-o", tmp]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
