"""Builds the oracle (test infrastructure, never part of the product).

* oracle/_build/liboracle32.so, liboracle64.so: the C restatement in
  oracle.c, FP32 bit-replay and FP64 builds (gcc, IEEE semantics, no FP
  contraction).
* oracle/_ref/libref_boundary.so: the reference's own boundary-physics header
  (src/optical/optix/OptiXBoundaryPhysics.cuh) compiled from where it lies in
  the reference tree, wrapped by ref_boundary_harness.cpp.  Built only when
  the reference tree is present; the built library is git-ignored and travels
  with the working tree.

The rest of the reference path (OptiXTransportDevice.cu, the OptiX host) needs
nvcc, OptiX and Geant4 and is not buildable here.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

ORACLE_DIR = Path(__file__).resolve().parent
BUILD_DIR = ORACLE_DIR / "_build"
REF_DIR = ORACLE_DIR / "_ref"
REFERENCE_ROOT = Path(os.environ.get("G4GO_REFERENCE_ROOT", "/root/reference"))
REFERENCE_OPTIX_SRC = REFERENCE_ROOT / "src" / "optical" / "optix"

CFLAGS = ["-O2", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wextra",
          "-Wno-unused-parameter"]

LIB32 = BUILD_DIR / "liboracle32.so"
LIB64 = BUILD_DIR / "liboracle64.so"
LIBREF = REF_DIR / "libref_boundary.so"


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{proc.stdout}")


def build_oracle() -> None:
    BUILD_DIR.mkdir(exist_ok=True)
    src = str(ORACLE_DIR / "oracle.c")
    _run(["gcc", *CFLAGS, src, "-o", str(LIB32), "-lm"])
    _run(["gcc", *CFLAGS, "-DORACLE_DOUBLE", src, "-o", str(LIB64), "-lm"])


def build_reference() -> bool:
    """Compiles the reference boundary-physics header; False if the tree is absent."""
    header = REFERENCE_OPTIX_SRC / "OptiXBoundaryPhysics.cuh"
    if not header.is_file():
        return False
    REF_DIR.mkdir(exist_ok=True)
    _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++20", "-ffp-contract=off", "-fno-fast-math",
          f"-I{REFERENCE_OPTIX_SRC}", str(ORACLE_DIR / "ref_boundary_harness.cpp"), "-o", str(LIBREF)])
    return True


def build() -> None:
    build_oracle()
    build_reference()


if __name__ == "__main__":
    build()
    print("oracle built:", LIB32, LIB64, LIBREF if LIBREF.exists() else "(reference absent)")
