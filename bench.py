#!/usr/bin/env python3
"""Benchmark of the MI355X optical-photon transport path.

Workload (one step = one batch through the hot path): 1e8 synthetic photons in
the 3x3x20 mm wrapped crystal (n = 1.82, absorption 400 mm, Rayleigh 500 mm,
R_wrap = 0.95, SiPM PDE curve on +z), positions uniform in the crystal,
isotropic, random linear polarisation, wavelength N(420, 20) nm, emission time
Exp(40 ns); outputs the compacted detected-hit list in HBM plus the 256-bin
0-200 ns arrival-time histogram.  Photons are generated on the device from
(seed, event, photon), so every input is resident before the timed region.

Timed: K steps of enqueue + device completion, bracketed by device
synchronisation.  `value` is photons/s of the persistent (batched) form; the
plain one-wave-per-64-photons launch is timed beside it.  The roofline object
uses the transport kernel's own HIP-event duration, averaged over the timed
steps, and the VALU lane-op model of the path (per generated photon, per
transport step, per detection).  The CPU baseline transports a bounded sample
of the same workload with the FP32 oracle on the host's cores.  The
`adjoint` object times a gradient pass of the differentiation workload (1e7
photons, 66 optical parameters, MSE loss on count + histogram; forward, loss
and the path-replay adjoint through torch.autograd) against a forward pass;
`array_events` times 1024 uneven electron events in the 8x8 crystal array
(batched vs per-event launches, 2048-parameter adjoint); `electron_events`
times 5000 1 MeV electron events in the crystal, with the FP32 oracle at 6
threads and at all host threads on a sample of them; `host_batches` times the
drop-in path at the reference scheduler's batch size (~1e6 photons, hits read
back to the host) with one and with two batches in flight; `mesh_crystals`
times 1e8 photons in two crystals with mesh solids: the crystal tapered to a
4x4 mm back face (a frustum mesh, source uniform in the 3x3x20 mm box inside
it) and the 3x3x20 mm crystal turned by a rotation (box meshes, point source
at its centre), both convex, a ring of 48 bare crystals turned to face its
centre (96 box meshes under a BVH; photons from crystal 0), and an L-shaped
light guide (a non-convex prism held as triangles), each with its
transport-step cost relative to the box crystal.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# VALU lane-op model (per unit; transcendental = 4): generated photon, step, detection.
W_GEN, W_STEP, W_DET = 290.0, 193.0, 15.0
VALU_PEAK_TLANEOPS = 39.3   # 256 CU x 64 lanes x 2.4 GHz
HBM_PEAK_GBS = 8000.0
HIT_BYTES = 48
PLAIN = dict(persistent=False, block_size=256, chunk=64, refill_threshold=64, blocks_per_cu=8)
TUNE_PHOTONS = 20_000_000


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--photons", type=int, default=100_000_000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--tune-table-dir", default=None,
                    help="directory that receives the autotuner's cache (and the crystal sweep table when it sweeps)")
    ap.add_argument("--cpu-sample", type=int, default=40_000_000, help="photons in the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-array", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-mesh", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the 1e8-photon batch: under a kernel trace, the transport kernel's average is that launch")
    ap.add_argument("--retune", action="store_true", help="ignore the autotuner cache")
    ap.add_argument("--launch", type=str, default=None,
                    help="explicit persistent form block,blocks_per_cu,chunk,refill (skips the autotuner)")
    return ap.parse_args()


def progress(msg: str) -> None:
    """A line on stderr as each part of the bench ends (stdout holds only the result)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def run_form(host, cs, src, photons, launch, steps, warmup):
    from g4go_amd.transport import LaunchForm

    host.set_launch(launch if isinstance(launch, LaunchForm) else LaunchForm(**launch))
    ev = np.array([0], np.uint32)
    cnt = np.array([photons], np.uint32)
    for _ in range(warmup):
        host.enqueue_synthetic(cs.scene, src, ev, cnt)
        host.complete_device()
    kernel_ms, outs = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        host.enqueue_synthetic(cs.scene, src, ev, cnt)
        out, _ = host.complete_device()
        kernel_ms.append(out.statistics["kernel_ms"])
        outs.append(out)
    wall = time.perf_counter() - t0
    return wall, kernel_ms, outs[-1]


def run_uneven(host, cs, src, counts, launch, steps, warmup):
    """Batches of uneven events: wall time per batch, kernel ms."""
    host.set_launch(launch)
    ev = np.arange(counts.size, dtype=np.uint32)
    for _ in range(warmup):
        host.enqueue_synthetic(cs.scene, src, ev, counts)
        host.complete_device()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        host.enqueue_synthetic(cs.scene, src, ev, counts)
        out, _ = host.complete_device()
        kms.append(out.statistics["kernel_ms"])
    wall = time.perf_counter() - t0
    return wall, float(np.mean(kms)), out


def run_adjoint(cs, src, seed, launch, steps, warmup, photons=10_000_000):
    """Differentiation workload: 1e7 photons, 66 parameters (32 absorption nodes,
    Rayleigh length, wrap reflectivity, 32 efficiency nodes), loss = (N - N*)^2 +
    sum_b (h_b - h*_b)^2 against a target run at a seed-perturbed (+-20%) theta.
    A gradient pass is forward + loss + torch.autograd backward (the path-replay
    adjoint kernel); it is timed against a forward pass on the same batch."""
    import torch

    from g4go_amd.autograd import histogram_loss, transport_histogram
    from g4go_amd.calibrate import clamp_probabilities, crystal_layout, perturbed
    from g4go_amd.transport import HipTransportHost, PhotonTransportConfig

    ev = np.array([0], np.uint32)
    cnt = np.array([photons], np.uint32)
    hosts = []
    for s in (seed, seed + 1000):
        h = HipTransportHost(PhotonTransportConfig(seed=s, sort_hits=False))
        h.set_launch(launch)
        h.prepare_scene(cs.scene)
        h.set_parameter_layout(crystal_layout(cs))
        hosts.append(h)
    host, target_host = hosts
    theta0 = host.get_parameters()
    with torch.no_grad():
        n_t, h_t = transport_histogram(torch.from_numpy(clamp_probabilities(perturbed(theta0, seed + 1), 32, 32)),
                                       target_host, cs.scene, src, ev, cnt)

    def forward():
        host.set_parameters(theta0)
        host.enqueue_synthetic(cs.scene, src, ev, cnt)
        out, _ = host.complete_device()
        return out.statistics["kernel_ms"]

    def gradient():
        theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
        n, hist = transport_histogram(theta, host, cs.scene, src, ev, cnt)
        histogram_loss(n, hist, n_t, h_t).backward()
        return theta.grad

    for _ in range(warmup):
        forward()
        gradient()
    t0 = time.perf_counter()
    f_kms = [forward() for _ in range(steps)]
    f_wall = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        g = gradient()
    g_wall = (time.perf_counter() - t0) / steps
    f_k = forward()
    _, stats = host.adjoint(np.ones(258))
    for h in hosts:
        h.close()
    return {"photons": photons, "parameters": int(theta0.size), "forward_ms": 1e3 * f_wall,
            "gradient_pass_ms": 1e3 * g_wall, "gradient_over_forward": g_wall / f_wall,
            "forward_kernel_ms": float(np.mean(f_kms)), "adjoint_kernel_ms": stats["kernel_ms"],
            "adjoint_over_forward_kernel": stats["kernel_ms"] / f_k, "replayed_photons": stats["replayed"],
            "replay_mismatches": stats["mismatched"], "grad_norm": float(torch.linalg.vector_norm(g))}


def run_emission_batches(host, scene, em, ev, idx, launch, steps, warmup):
    """Emission-record batches: wall seconds per batch, mean kernel ms, last output."""
    host.set_launch(launch)
    for _ in range(warmup):
        host.enqueue_emissions(scene, em, ev, idx)
        host.complete_device()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        host.enqueue_emissions(scene, em, ev, idx)
        out, _ = host.complete_device()
        kms.append(out.statistics["kernel_ms"])
    return (time.perf_counter() - t0) / steps, float(np.mean(kms)), out


def run_array(seed, steps, warmup, plain, retune=False):
    """8x8 array with septa (128 boxes), 1024 uneven electron events (log-normal
    median 2e4 photons): batched (persistent + hipGraph, launch form autotuned
    on this batch and cached under its own tag) vs per-event plain launches,
    and the 2048-parameter per-crystal absorption gradient."""
    from g4go_amd import abi, autotune, workloads
    from g4go_amd.transport import HipTransportHost, LaunchForm, PhotonTransportConfig

    cs = workloads.array_scene()
    em, ev, idx, _ = workloads.array_events(cs, 1024, median=2e4, seed=seed)
    photons = int(em["photon_count"].astype(np.int64).sum())
    host = HipTransportHost(PhotonTransportConfig(seed=seed, sort_hits=False))
    host.prepare_scene(cs.scene)
    host.set_parameter_layout([(abi.PARAM_ABS_LENGTH, m) for m in cs.crystal_materials])

    def tune_batch():
        host.enqueue_emissions(cs.scene, em, ev, idx)
        out, _ = host.complete_device()
        return out.statistics["kernel_ms"]

    launch, _ = autotune.tuned_launch(host, autotune.device_key(host), "array-emission", tune_batch, retune=retune)
    graphed = LaunchForm(**{**launch.__dict__, "graph": True})
    b_wall, b_kms, b_out = run_emission_batches(host, cs.scene, em, ev, idx, graphed, steps, warmup)
    _, adj = host.adjoint(np.ones(abi.HIST_SLOTS))
    res = {"events": 1024, "photons": photons, "boxes": 128, "parameters": 64 * 32,
           "batched_photons_per_s": photons / b_wall, "batched_kernel_ms": b_kms,
           "steps_per_photon": b_out.statistics["total_steps"] / max(1, b_out.statistics["generated"]),
           "detected_fraction": b_out.statistics["detected"] / max(1, b_out.statistics["generated"]),
           "adjoint_kernel_ms": adj["kernel_ms"], "adjoint_over_forward_kernel": adj["kernel_ms"] / b_kms,
           "replay_mismatches": adj["mismatched"], "launch": launch.__dict__}
    if plain:
        p_wall, p_kms, p_out = run_emission_batches(host, cs.scene, em, ev, idx, LaunchForm(persistent=False),
                                                    steps, warmup)
        e_wall, e_kms, e_out = run_emission_batches(host, cs.scene, em, ev, idx,
                                                    LaunchForm(persistent=False, per_event=True), steps, warmup)
        assert e_out.statistics["detected"] == b_out.statistics["detected"] == p_out.statistics["detected"]
        res.update({"plain_batch_photons_per_s": photons / p_wall, "plain_batch_kernel_ms": p_kms,
                    "batched_over_plain_batch": p_wall / b_wall,
                    "plain_per_event_photons_per_s": photons / e_wall, "plain_per_event_kernel_ms": e_kms,
                    "batched_over_plain_per_event": e_wall / b_wall})
    host.close()
    return res


def run_mesh(seed, steps, warmup, photons, box_ms_per_step, retune=False):
    """1e8 photons in the tapered crystal, the rotated crystal, the ring of 48
    turned crystals and the L-shaped light guide (mesh solids), each in the
    launch form autotuned on a 2e7-photon batch of it (cached per workload)."""
    from g4go_amd import autotune, workloads
    from g4go_amd.transport import HipTransportHost, PhotonTransportConfig

    rot = workloads.rotation_matrix((1.0, 2.0, 3.0), 0.7)
    tapered = workloads.tapered_crystal_scene()
    rotated = workloads.rotated_crystal_scene(rot)
    guide, guide_src, _, _ = workloads.l_guide_scene()
    guide_cs = workloads.CrystalScene(guide, guide.world_volume_id, [], [], [], -1, -1, workloads.CrystalParams())
    ring, ring_src = workloads.ring_scene()
    cases = {"tapered_crystal": (tapered, workloads.synthetic_source(tapered)),
             "rotated_crystal": (rotated, workloads.point_source(rotated.crystal_ids[0], (0.0, 0.0, 0.0))),
             "crystal_ring_48": (ring, ring_src),
             "l_guide_triangles": (guide_cs, guide_src)}
    res = {}
    for name, (cs, src) in cases.items():
        host = HipTransportHost(PhotonTransportConfig(seed=seed, sort_hits=False))
        one = np.array([0], np.uint32)

        def tune_batch():
            host.enqueue_synthetic(cs.scene, src, one, np.array([20_000_000], np.uint32))
            o, _ = host.complete_device()
            return o.statistics["kernel_ms"]

        launch, _ = autotune.tuned_launch(host, autotune.device_key(host), "mesh-" + name, tune_batch, retune=retune)
        wall, kms, out = run_form(host, cs, src, photons, launch, steps, warmup)
        host.close()
        st = out.statistics
        k_ms = float(np.mean(kms))
        res[name] = {"photons_per_s": photons * steps / wall, "kernel_ms": k_ms,
                     "steps_per_photon": st["total_steps"] / max(1, st["generated"]),
                     "detected_fraction": st["detected"] / max(1, st["generated"]),
                     "invalid": st["invalid"],
                     "step_cost_over_box": (k_ms / st["total_steps"]) / box_ms_per_step,
                     "launch": launch.__dict__}
    return res


def run_c1(seed, steps, warmup, cpu, cpu_events, retune=False):
    """5000 events of 1 MeV electrons in the crystal (scintillation records, 30
    photons/keV): device photons/s (launch form autotuned on this batch and
    cached under its own tag), and the FP32 oracle on `cpu_events` events at 6
    threads and at all (up to 16) host threads."""
    from g4go_amd import autotune, workloads
    from g4go_amd.transport import HipTransportHost, PhotonTransportConfig

    cs = workloads.crystal_scene()
    em, ev, idx = workloads.electron_emissions(cs, 5000, seed=seed)
    photons = int(em["photon_count"].astype(np.int64).sum())
    host = HipTransportHost(PhotonTransportConfig(seed=seed, sort_hits=False))
    host.prepare_scene(cs.scene)

    def tune_batch():
        host.enqueue_emissions(cs.scene, em, ev, idx)
        o, _ = host.complete_device()
        return o.statistics["kernel_ms"]

    launch, _ = autotune.tuned_launch(host, autotune.device_key(host), "crystal-electrons", tune_batch, retune=retune)
    wall, kms, out = run_emission_batches(host, cs.scene, em, ev, idx, launch, steps, warmup)
    host.close()
    res = {"events": 5000, "photons": photons, "photons_per_s": photons / wall, "kernel_ms": kms,
           "h2d_ms": out.statistics["host_to_device_ms"], "record_bytes": int(em.nbytes),
           "launch": launch.__dict__}
    if cpu:
        from oracle import oracle as O

        sel = idx < cpu_events
        n = int(em["photon_count"][sel].astype(np.int64).sum())
        for threads in (6, max(1, min(16, os.cpu_count() or 1))):
            t0 = time.perf_counter()
            O.run_emissions(cs.scene, em[sel], ev[:cpu_events], idx[sel], seed, threads=threads)
            dt = time.perf_counter() - t0
            res[f"cpu_{threads}_threads_photons_per_s"] = n / dt
        res["cpu_sample"] = f"first {cpu_events} events ({n} photons), FP32 oracle (oracle/oracle.c)"
    return res


def run_dropin_cpp(batches=200):
    """The drop-in path through the C++ shim (tests/cpp/shim_test bench): batches
    of 33 electron events (~1e6 photons) through PrepareScene / EnqueueEmissions /
    CompleteOldestBatch, hits returned as PhotonDetection records in launch order
    (converted from the library's hit records by eight host threads), with one
    batch in flight (the reference's single slot) and with two."""
    import subprocess

    exe = os.path.join(ROOT, "tests", "cpp", "shim_test")
    if not os.path.isfile(exe):
        return None
    res = {}
    for slots in (1, 2):
        proc = subprocess.run([exe, "bench", str(slots), str(batches), "8"], capture_output=True, text=True, timeout=120)
        if proc.returncode != 0:
            return {"error": proc.stdout[-500:] + proc.stderr[-500:]}
        res[f"slots_{slots}"] = json.loads(proc.stdout.strip().splitlines()[-1])
    return res


def run_host_batches(seed, n_batches=40, events_per_batch=33):
    """The drop-in path at the reference scheduler's batch size (~1e6 photons,
    PhotonTransportConfig::targetPhotonsPerBatch): 1 MeV electron events in the
    crystal, enqueue + completion with the hits read back to the host in launch
    order, with one batch in flight (the reference's single slot) and with two.
    The launch form is a new handle's default (what a drop-in caller gets, as
    in the C++ shim bench), replayed as a graph."""
    from g4go_amd import workloads
    from g4go_amd.transport import HipTransportHost, LaunchForm, PhotonTransportConfig

    cs = workloads.crystal_scene()
    em, ev, idx = workloads.electron_emissions(cs, n_batches * events_per_batch, seed=seed)
    batches = []
    for b in range(n_batches):
        sel = (idx >= events_per_batch * b) & (idx < events_per_batch * (b + 1))
        batches.append((em[sel], ev[events_per_batch * b:events_per_batch * (b + 1)],
                        (idx[sel] - events_per_batch * b).astype(np.uint32)))
    photons = int(em["photon_count"].astype(np.int64).sum())
    res = {"batches": n_batches, "photons_per_batch": photons / n_batches, "readback": "host, launch order"}
    for slots in (1, 2):
        host = HipTransportHost(PhotonTransportConfig(seed=seed, max_in_flight=slots))
        host.set_launch(LaunchForm(graph=True))
        host.prepare_scene(cs.scene)
        for rep in range(2):  # the first pass warms the graphs and buffers
            t0 = time.perf_counter()
            inflight = 0
            for e, i, x in batches:
                if inflight == slots:
                    host.complete_oldest_batch()
                    inflight -= 1
                host.enqueue_emissions(cs.scene, e, i, x)
                inflight += 1
            while inflight:
                host.complete_oldest_batch()
                inflight -= 1
            dt = time.perf_counter() - t0
        res[f"slots_{slots}_photons_per_s"] = photons / dt
        host.close()
    return res


# Host cores one GPU's job may use on the GPU host (its CPU share); the
# oracle's thread pool is sized to it, whatever the machine's core count is.
HOST_CORE_SHARE = 16


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cs, src, seed, sample):
    """The FP32 oracle on the host: the first `sample` photons of the headline
    workload at all the cores of the job's host share and at 6 threads (the
    reference's Geant4 CPU backend configuration)."""
    from oracle import oracle as O

    cores = max(1, min(HOST_CORE_SHARE, os.cpu_count() or 1))
    rates = {}
    for threads in (cores, 6):
        n = sample if threads == cores else sample * 6 // cores
        t0 = time.perf_counter()
        O.run_synthetic(cs.scene, src, seed, [0], [n], precision=32, threads=threads)
        rates[threads] = (n, time.perf_counter() - t0)
    n, dt = rates[cores]
    n6, dt6 = rates[6]
    return {"value": n / dt, "unit": "photons/s", "cores": cores, "kind": "port",
            "sample": f"photons 0..{n - 1} of the same 1e8-photon workload, FP32 oracle (oracle/oracle.c), "
                      f"{cores} threads, {dt:.2f} s wall",
            "value_6_threads": n6 / dt6, "sample_6_threads": f"photons 0..{n6 - 1}, 6 threads, {dt6:.2f} s wall",
            "cpu_model": cpu_model(), "host_cpus_visible": os.cpu_count(),
            "host_core_share": HOST_CORE_SHARE}


def profiled_counters(photons):
    """Calibrated rocprofv3 counters of this workload (profiles/), or None."""
    path = os.path.join(ROOT, "profiles", "transport_kernel_crystal_1e8_counters.json")
    if not os.path.isfile(path):
        return None
    with open(path) as f:
        data = json.load(f)
    return data if data.get("photons") == photons else None


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this backend runs on one GPU; multi-GPU runs are not part of this build "
                                   f"(asked for --gpus {args.gpus})"}))
        sys.exit(2)
    import torch  # device enumeration only; the transport runs through libg4go_hip.so

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    from g4go_amd import autotune, workloads
    from g4go_amd.transport import HipTransportHost, LaunchForm, PhotonTransportConfig

    cs = workloads.crystal_scene()
    src = workloads.synthetic_source(cs)
    host = HipTransportHost(PhotonTransportConfig(seed=args.seed, sort_hits=False))
    device, _ = host.device_info()
    tune_key = autotune.device_key(host)
    ev1 = np.array([0], np.uint32)

    def tune_batch():
        host.enqueue_synthetic(cs.scene, src, ev1, np.array([TUNE_PHOTONS], np.uint32))
        out, _ = host.complete_device()
        return out.statistics["kernel_ms"]

    if args.launch:
        bs, bpc, chunk, refill = (int(x) for x in args.launch.split(","))
        persistent = LaunchForm(persistent=True, block_size=bs, chunk=chunk, refill_threshold=refill,
                                blocks_per_cu=bpc)
    else:
        persistent, table = autotune.tuned_launch(host, tune_key, "crystal-synthetic", tune_batch,
                                                  retune=args.retune)
        out_dir = args.tune_table_dir
        if table is not None and out_dir and os.path.isdir(out_dir):
            with open(os.path.join(out_dir, "autotune_table.json"), "w") as f:
                json.dump(table, f, indent=0)
    wall, kms, out = run_form(host, cs, src, args.photons, persistent, args.steps, args.warmup)
    progress(f"headline: {args.photons * args.steps / wall:.4g} photons/s, kernel {float(np.mean(kms)):.3f} ms")
    if args.headline_only:
        args.no_plain = args.no_adjoint = args.no_array = args.no_c1 = args.no_mesh = True
    plain = None
    if not args.no_plain:
        p_wall, p_kms, p_out = run_form(host, cs, src, args.photons, PLAIN, args.steps, args.warmup)
        plain = {"photons_per_s": args.photons * args.steps / p_wall, "ms_per_step": 1e3 * p_wall / args.steps,
                 "kernel_ms": float(np.mean(p_kms))}
        progress(f"plain: kernel {plain['kernel_ms']:.3f} ms")
    # uneven events: 4096 log-normal events per batch, batched (persistent + hipGraph) vs plain per event
    counts = workloads.lognormal_counts(4096, 5000, sigma=1.5, lo=10, hi=1_000_000, seed=args.seed)
    n_uneven = int(counts.astype(np.int64).sum())
    uneven = None
    if not args.headline_only:
        ev_ids = np.arange(counts.size, dtype=np.uint32)

        def tune_uneven():
            host.enqueue_synthetic(cs.scene, src, ev_ids, counts)
            o, _ = host.complete_device()
            return o.statistics["kernel_ms"]

        uneven_form, _ = autotune.tuned_launch(host, tune_key, "crystal-uneven", tune_uneven, retune=args.retune)
        graphed = LaunchForm(**{**uneven_form.__dict__, "graph": True})
        b_wall, b_kms, b_out = run_uneven(host, cs, src, counts, graphed, args.steps, args.warmup)
        uneven = {"events": int(counts.size), "photons": n_uneven, "median": 5000, "sigma": 1.5,
                  "batched_photons_per_s": n_uneven * args.steps / b_wall, "batched_kernel_ms": b_kms,
                  "batched_launch": graphed.__dict__}
    if not args.no_plain:
        # the plain form as one launch over the whole batch (one wave per 64
        # photons of an event), and as one launch per event
        p_wall, p_kms, p_out = run_uneven(host, cs, src, counts, LaunchForm(persistent=False), args.steps,
                                          args.warmup)
        pe = LaunchForm(persistent=False, per_event=True)
        e_wall, e_kms, e_out = run_uneven(host, cs, src, counts, pe, args.steps, args.warmup)
        assert e_out.statistics["detected"] == b_out.statistics["detected"] == p_out.statistics["detected"]
        uneven.update({"plain_batch_photons_per_s": n_uneven * args.steps / p_wall,
                       "plain_batch_kernel_ms": p_kms,
                       "batched_over_plain_batch": p_wall / b_wall,
                       "plain_per_event_photons_per_s": n_uneven * args.steps / e_wall,
                       "plain_per_event_kernel_ms": e_kms,
                       "batched_over_plain_per_event": e_wall / b_wall})
    progress("uneven events done")
    host.close()
    adjoint = None if args.no_adjoint else run_adjoint(cs, src, args.seed, persistent, args.steps, args.warmup)
    progress("adjoint done")
    array = None if args.no_array else run_array(args.seed, args.steps, args.warmup, not args.no_plain,
                                                     retune=args.retune)
    progress("array done")
    c1 = None if args.no_c1 else run_c1(args.seed, args.steps, args.warmup, not args.no_cpu, 200,
                                               retune=args.retune)
    progress("electron events done")
    host_batches = None if args.no_c1 else run_host_batches(args.seed)
    dropin_cpp = None if args.no_c1 else run_dropin_cpp()
    progress("drop-in batches done")
    mesh = None if args.no_mesh else run_mesh(args.seed, args.steps, args.warmup, args.photons,
                                              float(np.mean(kms)) / out.statistics["total_steps"],
                                              retune=args.retune)
    progress("mesh crystals done")

    st = out.statistics
    P, S, D = st["generated"], st["total_steps"], st["n_hits"]
    k_ms = float(np.mean(kms))
    work = W_GEN * P + W_STEP * S + W_DET * D
    achieved = work / (k_ms * 1e-3) / 1e12
    alg_bytes = HIT_BYTES * D + 8 * 258 * 2
    prof = profiled_counters(args.photons)
    traffic = None if prof is None else prof["hbm_bytes_per_launch"]
    result = {
        "metric": "optical photons/sec on 1 MI355X (batched vs plain) + adjoint/forward cost ratio",
        "value": args.photons * args.steps / wall,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "1e8 synthetic photons, 3x3x20 mm wrapped crystal + SiPM, one batch, hits + 256-bin "
                        "0-200 ns histogram in HBM",
            "photons": args.photons, "form": "persistent", "launch": persistent.__dict__, "seed": args.seed,
            "device": device,
            "time": "fp64",
        },
        "plain": plain,
        "uneven_events": uneven,
        "adjoint": adjoint,
        "array_events": array,
        "electron_events": c1,
        "host_batches": host_batches,
        "dropin_cpp": dropin_cpp,
        "mesh_crystals": mesh,
        "batched_over_plain": (None if plain is None else (args.photons * args.steps / wall) / plain["photons_per_s"]),
        "counts": {"photons": P, "steps": S, "detected": D, "steps_per_photon": S / max(P, 1),
                   "detected_fraction": D / max(P, 1)},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TLANEOPS, "unit": "Tlane-op/s",
                     "frac": achieved / VALU_PEAK_TLANEOPS, "traffic": traffic,
                     "kernel": "transport_kernel_persistent<synthetic>", "kernel_ms": k_ms,
                     "work_model": {"per_photon": W_GEN, "per_step": W_STEP, "per_detection": W_DET}},
        "valu_counters": None if prof is None else {
            "source": "profiles/transport_kernel_crystal_1e8_counters.md",
            "VALUBusy_pct": prof["valu"]["VALUBusy"], "VALUUtilization_pct": prof["valu"]["VALUUtilization"],
            "active_lane_ops_per_step": prof["sq"]["SQ_INSTS_VALU"] * 64 * prof["valu"]["VALUUtilization"] / 100
            / prof["steps"]},
        "roofline_hbm": {"achieved": alg_bytes / (k_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": alg_bytes / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "bytes_per_launch": alg_bytes},
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(cs, src, args.seed, args.cpu_sample)
    if args.tune_table_dir and os.path.isdir(args.tune_table_dir) and autotune.CACHE.is_file():
        shutil.copy(autotune.CACHE, os.path.join(args.tune_table_dir, "autotune_cache.json"))
    print(json.dumps(result))


if __name__ == "__main__":
    main()
