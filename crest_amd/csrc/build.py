"""Build libcrest_hip.so (gfx950) in-tree with hipcc: one object per source, then link.

Usage: python crest_amd/csrc/build.py [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG = HERE.parent
ROOT = PKG.parent
OUT = PKG / "libcrest_hip.so"
OBJ = HERE / "_build"
SOURCES = ["api.hip", "kernels_chunk.hip", "kernels_gram.hip", "kernels_knn.hip", "kernels_graph.hip", "kernels_eig.hip", "kernels_sytrd.hip", "kernels_hvg.hip",
           "kernels_sparse.hip", "kernels_chefsi.hip"]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

CFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # every rounding is explicit: no silent FMA contraction
    "-ffp-contract=off",
    # native FP64 atomics in LDS and global memory
    "-munsafe-fp-atomics",
    # packed FP32 arithmetic is kept off on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROCM}/include", f"-I{ROOT / 'include'}",
]


def _hipcc() -> str:
    return os.path.join(ROCM, "bin", "hipcc")


def _compile(src: str) -> Path:
    obj = OBJ / (Path(src).stem + ".o")
    cmd = [_hipcc(), *CFLAGS, "-c", str(HERE / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def build(jobs: int = 4) -> Path:
    OBJ.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp),
           f"-L{ROCM}/lib", "-lrocsolver", "-lrocblas", f"-Wl,-rpath,{ROCM}/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libcrest_hip.so failed")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    print(build(ap.parse_args().jobs))
