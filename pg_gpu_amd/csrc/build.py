"""Build libpgld.so (gfx950) in-tree with hipcc.

    python pg_gpu_amd/csrc/build.py [--jobs N]

Regenerates the statistic-set include files, compiles every .hip file of this
directory for gfx950 and links pg_gpu_amd/libpgld.so.  Packed FP32 VALU
instructions are disabled for all device code.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG = HERE.parent
OUT = PKG / "libpgld.so"
OBJ = HERE / "_build"

ARCH = os.environ.get("PGLD_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unknown-warning-option",
    # a wave-uniform atomic returns its value to one lane; the optimizer's
    # rewrite would make every lane wait for it at once
    "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libpgld.so)")


def compile_one(cc, src):
    obj = OBJ / (src.stem + ".o")
    cmd = [cc, *FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(jobs=None, verbose=False):
    subprocess.run([sys.executable, str(HERE / "gen_stats.py")], check=True)
    cc = hipcc()
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(HERE.glob("*.hip"))
    jobs = jobs or min(len(srcs), os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(lambda s: compile_one(cc, s), srcs))
    if verbose:
        for _, err in results:
            if err.strip():
                print(err)
    objs = [str(o) for o, _ in results]
    tmp = OUT.with_suffix(".so.tmp")
    r = subprocess.run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
