"""Build libgpk.so (gfx950) in-tree with hipcc.

Usage: python gpjax_amd/csrc/build.py [--jobs N] [--verbose]

Each .hip source compiles to an object under gpjax_amd/csrc/_build/ and the
objects link into gpjax_amd/libgpk.so.  Every invocation recompiles (sources
are few and hipcc is fast); objects are written to temporary names and the
library is replaced atomically.
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

CSRC = Path(__file__).resolve().parent
PKG = CSRC.parent
OUT = PKG / "libgpk.so"
BUILD = CSRC / "_build"
SOURCES = ["api.hip", "gram.hip", "leaf.hip", "gemm.hip", "adjoint.hip", "vec.hip", "elbo.hip",
           "fit.hip"]
ARCH = "gfx950"
# Packed fp32 ops are disabled for every device pass: some v_pk_*_f32 forms have
# produced silent wrong results on these parts.
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: libgpk needs the ROCm toolchain to build")


def _compile(src: str, verbose: bool) -> Path:
    obj = BUILD / (Path(src).stem + ".o")
    tmp = obj.with_suffix(".o.tmp")
    cmd = [hipcc(), *FLAGS, "-c", str(CSRC / src), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in proc.stderr.splitlines()
                     if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{errs}")
    if errs.strip() and verbose:
        print(errs, file=sys.stderr)
    tmp.replace(obj)
    return obj


def build(jobs: int = 6, verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = OUT.with_name(OUT.name + ".tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stderr}")
    tmp.replace(OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=6)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
