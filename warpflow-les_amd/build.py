"""Build the gfx950 HIP library ``warpflow-les_amd/lib/libwarpflow_hip.so``.

    python warpflow-les_amd/build.py [--jobs N] [--no-contract]

Each ``csrc/*.hip`` is compiled to an object with hipcc for gfx950 and the
objects are linked into one shared library exporting the C ABI declared in
``include/warpflow_hip.h``.  The library is built in-tree so it travels with
the repository snapshot to the GPU machine.
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
CSRC = HERE / "csrc"
OBJ = HERE / "_build"
LIB = HERE / "lib" / "libwarpflow_hip.so"
ARCH = "gfx950"

# -packed-fp32-ops: keep the compiler from emitting packed FP32 VALU ops.
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build warpflow)")


def _compile(src: Path, extra: list[str]) -> Path:
    obj = OBJ / (src.stem + ".o")
    cmd = [hipcc(), *COMMON, *extra, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src.name, r.stdout, r.stderr))
    return obj


def build(jobs: int = 4, contract: bool = True, verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    extra = [] if contract else ["-ffp-contract=off"]
    srcs = sorted(CSRC.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: _compile(s, extra), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs),
           "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--no-contract", action="store_true",
                    help="build with -ffp-contract=off (ulp-level parity diagnosis)")
    a = ap.parse_args(argv)
    build(a.jobs, contract=not a.no_contract, verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
