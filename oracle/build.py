"""Build the C oracle (test infrastructure) into oracle/_build/libwf_oracle.so.

    python oracle/build.py

Plain gcc, -O3 with OpenMP, strict IEEE FP64 (no -ffast-math, no FMA
contraction) so it agrees with the NumPy restatement to rounding.

The reference itself has no C/C++ sources for this path (it is Python +
NVIDIA Warp kernels JIT-compiled at run time), so there is no oracle/_ref
build: the path is unbuildable as C here and the oracle is pinned by the
checks in tests/test_oracle_pins.py instead.
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "c" / "wf_oracle.c"
OUT = HERE / "_build" / "libwf_oracle.so"


def build(verbose: bool = False) -> Path:
    OUT.parent.mkdir(exist_ok=True)
    cmd = ["gcc", "-O3", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           "-fno-fast-math", str(SRC), "-o", str(OUT), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("gcc failed:\n" + r.stdout + r.stderr)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
