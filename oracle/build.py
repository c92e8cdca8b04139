"""ORACLE build recipe (test infrastructure): compiles crest_oracle.c into
oracle/liboracle.so with gcc.  Floating-point contraction is off so every multiply
and add rounds separately, as the reference's Rust code does.

The reference itself (Rust + PyO3, crates not vendored, no cargo in this image) is
unbuildable here, so there is no oracle/_ref build.
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "liboracle.so"


def build() -> Path:
    tmp = OUT.with_suffix(".so.tmp")
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-shared",
           str(HERE / "crest_oracle.c"), "-o", str(tmp), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("gcc failed on oracle/crest_oracle.c")
    tmp.replace(OUT)
    return OUT


if __name__ == "__main__":
    print(build())
