"""Build recipe for the gfx950 library (libveig.so) and the CPU oracle.

    python vectorized-eigensolver_amd/build.py          # library + oracle
    python vectorized-eigensolver_amd/build.py --lib    # library only

The library is compiled in-tree with hipcc for gfx950 only.  Packed FP32
instructions are disabled (-packed-fp32-ops) for all device code.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libveig.so"
ORACLE = REPO / "oracle"

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libveig.so)")


def build_lib(verbose: bool = False) -> Path:
    srcs = [str(CSRC / "veig.hip"), str(CSRC / "veig_mixed.hip"), str(CSRC / "veig_compat.cpp")]
    srcs = [s for s in srcs if Path(s).exists()]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", str(REPO / "include"), "-o", str(tmp), *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode})")
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = False) -> None:
    targets = ["all"]
    if Path("/root/reference/utils.cpp").exists():
        targets.append("ref")
    cmd = ["make", "-C", str(ORACLE), "-j4", *targets]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("oracle build failed")


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("--oracle", action="store_true", help="oracle only")
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    if not a.oracle:
        print("built", build_lib(a.v))
    if not a.lib:
        build_oracle(a.v)
        print("built oracle")


if __name__ == "__main__":
    main()
