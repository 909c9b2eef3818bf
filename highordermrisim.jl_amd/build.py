"""Compiles libhom.so (HIP, gfx950) in-tree with hipcc.

Every HIP source is compiled to an object in parallel, then linked into
``libhom.so`` next to this file.  Flags: -O3, gfx950 only, and packed-FP32
instruction selection disabled (``-packed-fp32-ops``)."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libhom.so"
SOURCES = ["hom_abi.hip", "plain_kernels.hip", "mixed_kernels.hip", "serve.hip", "staging.cpp", "kspace.cpp", "discretize.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc" if Path("/opt/rocm/bin/hipcc").exists()
                       else "hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", str(HERE.parent / "include")]


def _compile(src: str, build_dir: Path) -> Path:
    obj = build_dir / (Path(src).stem + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    build_dir = HERE / "_build"
    build_dir.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(lambda s: _compile(s, build_dir), SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    shutil.rmtree(build_dir, ignore_errors=True)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
