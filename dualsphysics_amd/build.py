"""Builds the native libraries in-tree.

- ``dualsphysics_amd/libdsph.so``: the HIP kernels for gfx950, the C ABI of
  ``include/dsph_abi.h`` and the C++ host driver (hipcc).
- ``oracle/liboracle.so``: the CPU oracle (g++/OpenMP), test infrastructure.
- ``oracle/_ref/``: the reference CPU solver built from its own sources
  (only where ``/root/reference`` exists), see ``oracle/ref.mk``.

Usage: ``python -m dualsphysics_amd.build [--no-oracle] [--no-ref] [-j N]``
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "dualsphysics_amd"
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = ROOT / "_build"
LIB = PKG / "libdsph.so"
ORACLE_LIB = ROOT / "oracle" / "liboracle.so"
REFERENCE = Path("/root/reference/src/source")

ARCH = os.environ.get("DSPH_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # Packed FP32 instructions off for every HIP translation unit.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # Contract a*b+c into FMA only within one source expression, so two kernels
    # that evaluate the same pair expressions produce the same bits.
    "-ffp-contract=on",
    "-Wno-unused-result",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def _sources():
    hip = sorted(CSRC.glob("*.hip"))
    cpp = sorted((CSRC / "host").glob("*.cpp"))
    return hip, cpp


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(map(str, cmd)) + "\n" + r.stdout + r.stderr)
    return r


def build_native(jobs: int = 8, verbose: bool = False) -> Path:
    hip, cpp = _sources()
    BUILD.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for src in hip + cpp:
        obj = BUILD / (src.stem + (".hip.o" if src.suffix == ".hip" else ".o"))
        objs.append(obj)
        if src.suffix == ".hip":
            cmds.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        else:
            # Host sources include the kernels' headers: same flags as the HIP files.
            cmds.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for r in ex.map(_run, cmds):
            if verbose and r.stderr:
                print(r.stderr, file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    srcs = [ROOT / "oracle" / "dsph_oracle.cpp", ROOT / "oracle" / "dsph_oracle_sim.cpp"]
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    cxx = shutil.which("g++") or "g++"
    # No FMA contraction: the reference CPU build the fixtures come from
    # (oracle/ref.mk, IEEE flags, baseline x86-64) has none, so the oracle's
    # CPU semantics reproduce its forces bit for bit; no fast-math relaxation.
    _run([cxx, "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
          "-std=c++17", f"-I{INCLUDE}", *map(str, srcs), "-o", str(tmp)])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build_reference(jobs: int = 8) -> Path | None:
    """Builds the reference CPU solver from its sources (this container only)."""
    if not REFERENCE.is_dir():
        return None
    r = subprocess.run(["make", "-f", "oracle/ref.mk", f"-j{jobs}", "all", "fast"], cwd=ROOT,
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("reference build failed:\n" + r.stdout[-4000:] + r.stderr[-4000:])
    return ROOT / "oracle" / "_ref" / "DualSPHysics5.4CPU_ieee"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("--no-ref", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    print("built", build_native(a.j, a.v))
    if not a.no_oracle:
        print("built", build_oracle())
    if not a.no_ref:
        p = build_reference(a.j)
        print("built", p) if p else print("reference sources absent: skipped oracle/_ref")


if __name__ == "__main__":
    main()
