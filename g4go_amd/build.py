"""Builds libg4go_hip.so (gfx950) in-tree with hipcc.

The HIP sources are compiled with IEEE single-precision semantics so that the
transport arithmetic is reproducible on the host: no FP contraction,
correctly rounded division and square root, and packed FP32 instructions
disabled (the compiler would otherwise form v_pk_* operations by itself).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
CSRC = PACKAGE_DIR / "csrc"
INCLUDE = PACKAGE_DIR.parent / "include"
LIBRARY = PACKAGE_DIR / "libg4go_hip.so"
SHIM_TEST_SRC = PACKAGE_DIR.parent / "tests" / "cpp" / "shim_test.cpp"
SHIM_TEST = PACKAGE_DIR.parent / "tests" / "cpp" / "shim_test"

SOURCES = ["kernels.hip", "capi.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    f"-I{INCLUDE}",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def build(verbose: bool = True) -> Path:
    """Compiles every HIP source for gfx950 and links libg4go_hip.so."""
    objects = []
    build_dir = PACKAGE_DIR / "_build"
    build_dir.mkdir(exist_ok=True)
    for src in SOURCES:
        obj = build_dir / (Path(src).stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        objects.append(str(obj))
    tmp = LIBRARY.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objects, "-o", str(tmp)])
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY}")
    build_shim_test(verbose)
    return LIBRARY


PROFILE_LIBRARY = PACKAGE_DIR / "_prof" / "libg4go_hip.so"


def build_profile(verbose: bool = True) -> Path:
    """The region-profile counter build (-DG4GO_REGION_PROFILE, csrc/region_prof.h) in _prof/:
    the same kernels with cycle counters at region markers; never the product library."""
    prof_dir = PROFILE_LIBRARY.parent
    prof_dir.mkdir(exist_ok=True)
    objects = []
    for src in SOURCES:
        obj = prof_dir / (Path(src).stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-DG4GO_REGION_PROFILE", "-c", str(CSRC / src), "-o", str(obj)])
        objects.append(str(obj))
    tmp = PROFILE_LIBRARY.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objects, "-o", str(tmp)])
    os.replace(tmp, PROFILE_LIBRARY)
    if verbose:
        print(f"built {PROFILE_LIBRARY}")
    return PROFILE_LIBRARY


def build_shim_test(verbose: bool = True) -> Path:
    """Host-only C++20 test program of include/g4go_hip_host.hpp, linked to the library."""
    tmp = SHIM_TEST.with_name(SHIM_TEST.name + ".tmp")
    _run(["g++", "-std=c++20", "-O2", "-pthread", "-Wall", "-Wextra", f"-I{INCLUDE}", str(SHIM_TEST_SRC), f"-L{PACKAGE_DIR}",
          "-lg4go_hip", "-Wl,-rpath,$ORIGIN/../../g4go_amd", "-o", str(tmp)])
    os.replace(tmp, SHIM_TEST)
    if verbose:
        print(f"built {SHIM_TEST}")
    return SHIM_TEST


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "profile":
        build_profile()
    else:
        build()
