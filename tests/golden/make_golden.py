#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/.

* philox_kat.json: the published Random123 known-answer vectors of
  Philox4x32-10 (Salmon et al., SC'11; Random123 kat_vectors).
* reference_boundary_cases.json: the cases and expected values of the
  reference's unit test test/TestOpticalBoundaryPhysics.cxx:37-166 (Fresnel
  at normal incidence, Brewster, TIR, surface classification, detection,
  Lambertian, border/skin surface precedence), restated as data.
* ref_boundary_outputs.npz: inputs and outputs of the reference's own
  boundary-physics header (src/optical/optix/OptiXBoundaryPhysics.cuh),
  compiled by oracle/build.py into oracle/_ref/libref_boundary.so, on seeded
  random inputs.  Needs the reference tree; the npz is committed so the
  comparison runs without it.

Run from the repository root: python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import json
import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

PHILOX_KAT = [
    {"ctr": [0x00000000, 0x00000000, 0x00000000, 0x00000000], "key": [0x00000000, 0x00000000],
     "out": [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]},
    {"ctr": [0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff], "key": [0xffffffff, 0xffffffff],
     "out": [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]},
    {"ctr": [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], "key": [0xa4093822, 0x299f31d0],
     "out": [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]},
]

BREWSTER = math.atan(1.5)
TIR = 50.0 * math.pi / 180.0

BOUNDARY_CASES = {
    "fresnel": [
        {"name": "normal incidence 1.0->1.5", "dir": [0, 0, 1], "pol": [1, 0, 0], "normal": [0, 0, -1], "n1": 1.0,
         "n2": 1.5, "transmittance": 0.96, "tol": 1e-5, "tir": False, "reflected_dir_z": -1.0,
         "transmitted_dir_z": 1.0},
        {"name": "p-polarised Brewster 1.0->1.5", "dir": [math.sin(BREWSTER), 0, math.cos(BREWSTER)],
         "pol": "p", "normal": [0, 0, -1], "n1": 1.0, "n2": 1.5, "transmittance": 1.0, "tol": 1e-5},
        {"name": "TIR at 50 deg 1.5->1.0", "dir": [math.sin(TIR), 0, math.cos(TIR)], "pol": [0, 1, 0],
         "normal": [0, 0, -1], "n1": 1.5, "n2": 1.0, "transmittance": 0.0, "tol": 1e-6, "tir": True},
    ],
    # ClassifySurface(R, T, hasR, hasT, u) -> 0 absorb, 2 direct transmit, 3 interaction
    "classify": [
        [0.8, 0.1, 1, 1, 0.05, 3], [0.8, 0.1, 1, 1, 0.75, 3], [0.8, 0.1, 1, 1, 0.85, 2], [0.8, 0.1, 1, 1, 0.95, 0],
        [0.0, 0.0, 0, 0, 0.5, 3], [0.0, 0.4, 0, 1, 0.2, 2],
    ],
    # EvaluateAbsorption(efficiency, sensor, u) -> 1 detect, 0 absorb
    "absorption": [[0.3, 1, 0.2, 1], [0.3, 1, 0.5, 0]],
    "lambertian": {"normal": [0, 0, 1], "radial": 0.75, "azimuth": 0.25},
    # volumes 0 (parent, skin 0), 1 (daughter of 0, skin 1), 2 (daughter of 0, skin 2); border 0->1 = surface 3
    "precedence": {
        "without_border": [[0, 1, 1], [1, 0, 1], [1, 2, 1]],
        "with_border": [[0, 1, 3]],
    },
}


def reference_outputs(n: int = 1500, seed: int = 20240611):
    from oracle import oracle as O

    lib = O.reference_library()
    if lib is None:
        return None
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm[np.sum(d * nrm, axis=1) > 0] *= -1
    p = rng.normal(size=(n, 3))
    n1 = rng.uniform(1.0, 2.0, n)
    n2 = rng.uniform(1.0, 2.0, n)
    cases = np.concatenate([d, p, nrm, n1[:, None], n2[:, None]], axis=1).astype(np.float32)
    cases[:16, 6:9] = [0, 0, -1]
    cases[:8, 0:3] = [0, 0, 1]
    fres = np.zeros((n, 14), np.float32)
    proj = np.zeros((n, 3), np.float32)
    rdir = np.zeros((n, 3), np.float32)
    rpol = np.zeros((n, 3), np.float32)
    for i in range(n):
        c = np.ascontiguousarray(cases[i])
        lib.ref_compute_fresnel(c.ctypes.data, fres[i].ctypes.data)
        lib.ref_project_polarization(c[3:6].ctypes.data, c[0:3].ctypes.data, proj[i].ctypes.data)
        lib.ref_reflect_direction(c[0:3].ctypes.data, c[6:9].ctypes.data, rdir[i].ctypes.data)
        lib.ref_reflect_polarization(c[3:6].ctypes.data, c[6:9].ctypes.data, rdir[i].ctypes.data,
                                     rpol[i].ctypes.data)
    cls_in = np.stack([rng.uniform(0, 1, n), rng.uniform(0, 0.5, n), rng.integers(0, 2, n), rng.integers(0, 2, n),
                       rng.uniform(-0.1, 1.1, n)], axis=1).astype(np.float32)
    cls_out = np.array([lib.ref_classify_surface(float(a), float(b), int(c), int(e), float(u))
                        for a, b, c, e, u in cls_in], np.int32)
    return dict(fresnel_in=cases, fresnel_out=fres, project_out=proj, reflect_dir_out=rdir, reflect_pol_out=rpol,
                classify_in=cls_in, classify_out=cls_out)


def main() -> None:
    (HERE / "philox_kat.json").write_text(json.dumps(PHILOX_KAT, indent=1) + "\n")
    (HERE / "reference_boundary_cases.json").write_text(json.dumps(BOUNDARY_CASES, indent=1) + "\n")
    out = reference_outputs()
    if out is not None:
        np.savez_compressed(HERE / "ref_boundary_outputs.npz", **out)
        print("wrote ref_boundary_outputs.npz from the compiled reference header")
    else:
        print("reference library absent: ref_boundary_outputs.npz left unchanged")


if __name__ == "__main__":
    main()
