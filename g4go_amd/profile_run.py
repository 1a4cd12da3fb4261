"""Runs one launch of a named workload, for profilers (rocprofv3 passes).

    python -m g4go_amd.profile_run --workload crystal --photons 100000000
    python -m g4go_amd.profile_run --workload uneven
    python -m g4go_amd.profile_run --workload tapered   # mesh solids: tapered crystal
    python -m g4go_amd.profile_run --workload rotated   # mesh solids: crystal turned by a rotation
    python -m g4go_amd.profile_run --workload lguide    # triangle solid: L-shaped light guide
    python -m g4go_amd.profile_run --workload ring      # 48 turned crystals (convex meshes)
    python -m g4go_amd.profile_run --copy-bytes 1073741824   # calibration copy
    python -m g4go_amd.profile_run --workload crystal --regions   # region profile (counter build)

--regions loads the region-profile counter build (g4go_amd/_prof, made by
`python -m g4go_amd.build profile`) instead of the product library and
prints, per region of the transport kernel (csrc/region_prof.h), its share of
the wave cycles and its lane utilisation.

The launch form is the autotuner result cached for the workload on the
device (for the mesh workloads their own tags, as bench.py tunes them), else
the crystal's, else the default persistent form.  --copy-bytes additionally runs one device
to device copy of that many bytes (hipMemcpy through torch), whose FETCH_SIZE
and WRITE_SIZE calibrate the traffic counters against a known byte count.
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("crystal", "uneven", "array", "tapered", "rotated", "lguide", "ring",
                                              "none"),
                    default="crystal")
    ap.add_argument("--photons", type=int, default=100_000_000)
    ap.add_argument("--events", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--form", choices=("tuned", "plain"), default="tuned")
    ap.add_argument("--copy-bytes", type=int, default=0)
    ap.add_argument("--repeat", type=int, default=1)
    ap.add_argument("--regions", action="store_true")
    ap.add_argument("--launch", default="",
                    help="persistent form block,blocks_per_cu,chunk,refill[,pregenerate] instead of the cached one")
    args = ap.parse_args(argv)

    from . import abi
    if args.regions:
        from .build import PROFILE_LIBRARY

        abi.LIBRARY_PATH = PROFILE_LIBRARY
    from . import autotune, workloads
    from .transport import HipTransportHost, LaunchForm, PhotonTransportConfig

    if args.copy_bytes:
        import torch

        a = torch.empty(args.copy_bytes // 4, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.copy_bytes}), flush=True)
    if args.workload == "none":
        return 0
    if args.workload == "array":
        cs = workloads.array_scene()
        src = workloads.synthetic_source(cs, crystal_index=27)
    elif args.workload == "tapered":
        cs = workloads.tapered_crystal_scene()
        src = workloads.synthetic_source(cs)
    elif args.workload == "rotated":
        cs = workloads.rotated_crystal_scene(workloads.rotation_matrix((1.0, 2.0, 3.0), 0.7))
        src = workloads.point_source(cs.crystal_ids[0], (0.0, 0.0, 0.0))
    elif args.workload == "ring":
        cs, src = workloads.ring_scene()
    elif args.workload == "lguide":
        scene, src, _, _ = workloads.l_guide_scene()
        cs = workloads.CrystalScene(scene, scene.world_volume_id, [], [], [], -1, -1, workloads.CrystalParams())
    else:
        cs = workloads.crystal_scene()
        src = workloads.synthetic_source(cs)
    host = HipTransportHost(PhotonTransportConfig(seed=args.seed, sort_hits=False))
    # each workload in the launch form cached for it (the crystal's otherwise)
    tag = {"uneven": "crystal-uneven", "tapered": "mesh-tapered_crystal", "rotated": "mesh-rotated_crystal",
           "ring": "mesh-crystal_ring_48", "lguide": "mesh-l_guide_triangles"}.get(args.workload, "crystal-synthetic")
    key = autotune.device_key(host)
    form = (autotune.cached(key, tag) or autotune.cached(key, "crystal-synthetic")
            or LaunchForm(persistent=True, block_size=512, chunk=256, refill_threshold=4, blocks_per_cu=2))
    if args.launch:
        v = [int(x) for x in args.launch.split(",")]
        form = LaunchForm(persistent=True, block_size=v[0], blocks_per_cu=v[1], chunk=v[2], refill_threshold=v[3],
                          pregenerate=bool(v[4]) if len(v) > 4 else False)
    if args.form == "plain":
        form = LaunchForm(persistent=False)
    if args.workload == "uneven":
        counts = workloads.lognormal_counts(args.events, 5000, sigma=1.5, lo=10, hi=1_000_000, seed=args.seed)
        form = LaunchForm(**{**form.__dict__, "graph": form.persistent})
    else:
        counts = np.array([args.photons], np.uint32)
    host.set_launch(form)
    ev = np.arange(counts.size, dtype=np.uint32)
    lib = abi.load()
    for _ in range(args.repeat):
        if args.regions:
            lib.g4go_region_profile_clear()
        host.enqueue_synthetic(cs.scene, src, ev, counts)
        out, _ = host.complete_device()
        st = out.statistics
        line = {"workload": args.workload, "photons": int(counts.astype(np.int64).sum()),
                "kernel_ms": st["kernel_ms"], "steps": st["total_steps"], "detected": st["detected"],
                "launch": form.__dict__}
        if args.regions:
            line["regions"] = region_table(lib, st["total_steps"])
        print(json.dumps(line), flush=True)
    host.close()
    return 0


REGIONS = ("refill", "generate", "step", "compact", "geometry", "rayleigh", "surface", "absorb_detect")


def region_table(lib, steps: int) -> dict:
    """Per region: share of the wave cycles of the four top-level regions, wave entries per 1000
    lane-steps, and lane utilisation (active lanes / 64 per entry)."""
    import ctypes as C

    raw = (C.c_ulonglong * (3 * len(REGIONS)))()
    if lib.g4go_region_profile_read(raw) != 0:
        raise RuntimeError("region profile read failed")
    v = np.array(raw[:], dtype=np.float64).reshape(len(REGIONS), 3)
    top = v[[0, 2, 3], 0].sum()
    out = {}
    for k, name in enumerate(REGIONS):
        cycles, entries, lanes = v[k]
        out[name] = {"cycle_share": cycles / top if top else 0.0,
                     "entries_per_1000_steps": 1000.0 * entries / max(steps, 1),
                     "lane_utilisation": lanes / (64.0 * entries) if entries else 0.0,
                     "cycles_per_entry": cycles / entries if entries else 0.0}
    return out


if __name__ == "__main__":
    sys.exit(main())
