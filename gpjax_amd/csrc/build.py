"""Build libgpk.so (gfx950) in-tree with hipcc.

Usage: python gpjax_amd/csrc/build.py [--jobs N] [--verbose] [--define NAME=VALUE]
       [--out PATH] [--sanitize]

Each .hip source compiles to an object under gpjax_amd/csrc/_build/ and the
objects link into gpjax_amd/libgpk.so.  Every invocation recompiles (sources
are few and hipcc is fast); objects are written to temporary names and the
library is replaced atomically.

--sanitize instruments the host code (the C ABI's argument checks, workspace
layout and planning) with AddressSanitizer and UBSan; the device code is built
as usual.  Such a library needs the ASan runtime preloaded and is for host-side
tests only (tests/test_host_sanitizers.py).
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
           "fit.hip", "linalg.hip"]
ARCH = "gfx950"
# Packed fp32 ops are disabled for every device pass: some v_pk_*_f32 forms have
# produced silent wrong results on these parts.
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


# Host-only sanitizer flags: each -fsanitize= sits behind -Xarch_host, so the
# device passes are not instrumented.
SANITIZE = [
    "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
    "-Xarch_host", "-fno-omit-frame-pointer",
]


def asan_runtime() -> Path:
    """The shared ASan runtime a --sanitize library links against."""
    clang = Path(hipcc()).resolve().parent.parent / "llvm" / "bin" / "clang++"
    if not clang.exists():
        clang = Path("/opt/rocm/llvm/bin/clang++")
    out = subprocess.run([str(clang), "-print-file-name=libclang_rt.asan-x86_64.so"],
                         capture_output=True, text=True, check=True).stdout.strip()
    return Path(out)


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: libgpk needs the ROCm toolchain to build")


def _compile(src: str, verbose: bool, defines=(), build_dir: Path = BUILD,
             extra=()) -> Path:
    obj = build_dir / (Path(src).stem + ".o")
    tmp = obj.with_suffix(".o.tmp")
    cmd = [hipcc(), *FLAGS, *extra, *[f"-D{d}" for d in defines], "-c", str(CSRC / src), "-o",
           str(tmp)]
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


def build(jobs: int = 6, verbose: bool = False, defines=(), out: Path = OUT,
          sanitize: bool = False, build_dir: Path | None = None) -> Path:
    """Compile every source and link ``out``; ``defines`` (NAME=VALUE strings)
    select compile-time variants of the kernels for A/B builds; ``sanitize``
    instruments the host code (ASan + UBSan)."""
    if build_dir is None:
        tag = "_".join(defines).replace("=", "-") + ("_san" if sanitize else "")
        build_dir = BUILD / tag if tag else BUILD
    build_dir = Path(build_dir)
    build_dir.mkdir(parents=True, exist_ok=True)
    extra = SANITIZE if sanitize else []
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose, defines, build_dir, extra), SOURCES))
    out = Path(out)
    tmp = out.with_name(out.name + ".tmp")
    link = [*SANITIZE, "-shared-libsan"] if sanitize else []
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *link, *map(str, objs), "-o",
           str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stderr}")
    tmp.replace(out)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=6)
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--define", action="append", default=[], metavar="NAME=VALUE")
    ap.add_argument("--out", default=str(OUT))
    ap.add_argument("--sanitize", action="store_true",
                    help="host AddressSanitizer + UBSan build (give --out another name)")
    a = ap.parse_args()
    if a.sanitize and Path(a.out).resolve() == OUT.resolve():
        ap.error("--sanitize needs --out: the in-tree libgpk.so stays uninstrumented")
    print(build(a.jobs, a.verbose, tuple(a.define), Path(a.out), a.sanitize))
