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
    """Compile each translation unit to an object in parallel, then link."""
    from concurrent.futures import ThreadPoolExecutor

    srcs = [CSRC / "veig.hip", CSRC / "veig_mixed.hip", CSRC / "veig_varsize.hip",
            CSRC / "veig_large.hip", CSRC / "veig_roots.hip", CSRC / "veig_band.hip",
            CSRC / "veig_big.hip", CSRC / "veig_compat.cpp"]
    srcs = [s for s in srcs if s.exists()]
    objdir = REPO / "_build"
    objdir.mkdir(exist_ok=True)
    cc = hipcc()

    def compile_one(src: Path) -> Path:
        obj = objdir / (src.name + ".o")
        cmd = [cc, *[f for f in HIPCC_FLAGS if f != "-shared"], "-c", "-I", str(REPO / "include"),
               "-o", str(obj), str(src)]
        if src.suffix == ".cpp":
            cmd = [cc, "-O3", "-std=c++17", "-fPIC", "-c", "-I", str(REPO / "include"), "-o", str(obj),
                   str(src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"hipcc failed on {src.name} ({res.returncode})")
        return obj

    with ThreadPoolExecutor(max_workers=min(4, len(srcs))) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"link failed ({res.returncode})")
    os.replace(tmp, LIB)
    return LIB


def build_probes(verbose: bool = False) -> None:
    """The test probe of the scalar complex API (include/veig_complex.hpp):
    a host build (g++) and a gfx950 build (hipcc) of tests/probe/."""
    probe = REPO / "tests" / "probe"
    cmds = [
        ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", str(REPO / "include"),
         "-o", str(probe / "libcomplex_probe_host.so"), str(probe / "complex_probe_host.cpp")],
        [hipcc(), *HIPCC_FLAGS, "-I", str(REPO / "include"),
         "-o", str(probe / "libcomplex_probe.so"), str(probe / "complex_probe.hip")],
    ]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"probe build failed: {cmd[-1]}")


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
        build_probes(a.v)
        print("built oracle and test probes")


if __name__ == "__main__":
    main()
