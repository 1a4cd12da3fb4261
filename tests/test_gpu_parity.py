"""GPU parity: the HIP kernels against the FP32 oracle on the same Philox streams.

Integer, index and record outputs are compared bit for bit: the detected-hit
list (in launch order), per-event counters and the arrival-time
histogram.  At full size (1e8 photons) the checks are size-independent
invariants (every photon ends in exactly one category, histogram total =
detections = hit count).
"""
import sys

import numpy as np
import pytest

from g4go_amd import abi, workloads
from g4go_amd.transport import HipTransportHost, LaunchForm, PhotonTransportConfig
from oracle import oracle as O

pytestmark = pytest.mark.gpu

STAT_FIELDS = ("detected", "absorbed", "escaped", "truncated", "max_bounce", "invalid", "zero_step")


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def host(seed=42, **kw):
    kw.setdefault("sort_hits", True)
    return HipTransportHost(PhotonTransportConfig(seed=seed, **kw))


def assert_same_hits(gpu_hits, ref_hits):
    assert gpu_hits.shape == ref_hits.shape, (gpu_hits.shape, ref_hits.shape)
    if gpu_hits.tobytes() != ref_hits.tobytes():
        a = gpu_hits.view(np.uint8).reshape(-1, 48)
        b = ref_hits.view(np.uint8).reshape(-1, 48)
        bad = np.nonzero(np.any(a != b, axis=1))[0]
        raise AssertionError(f"{bad.size} hit records differ; first {bad[:5]}: gpu {gpu_hits[bad[0]]} "
                             f"oracle {ref_hits[bad[0]]}")


def assert_same_stats(out, ref):
    ev = out.event_statistics
    for k, name in enumerate(("detected", "absorbed", "escaped", "truncated", "max_bounce", "invalid", "zero_step")):
        np.testing.assert_array_equal(ev[name], ref.stats[:, k].astype(np.uint64), err_msg=name)
    np.testing.assert_array_equal(out.histogram, ref.hist)
    assert out.statistics["total_steps"] == int(ref.stats[:, 7].sum())
    assert out.statistics["generated"] == int(ref.stats[:, 8].sum())


def test_device_math_bit_exact():
    log("case: device log/exp/sincos vs oracle")
    rng = np.random.default_rng(1)
    words = rng.integers(0, 2**32, 4000, dtype=np.uint64)
    u = ((words.astype(np.float32) + np.float32(0.5)) / np.float32(4294967296.0)).astype(np.float32)
    xs_log = np.concatenate([u, np.float32([1.0, 0.5, 2.0, 1e-30, 3e-39, 1e30, 123.456])]).astype(np.float32)
    xs_exp = rng.uniform(-90, 90, 4000).astype(np.float32)
    xs_sc = np.concatenate([u, np.float32([0.0, 0.125, 0.25, 0.375, 0.5, 0.75, 0.999999])]).astype(np.float32)
    with host() as h:
        got_log = h.eval_math(0, xs_log)
        got_exp = h.eval_math(1, xs_exp)
        got_sc = h.eval_math(2, xs_sc)
    ref_log = np.array([O.log(float(x)) for x in xs_log], np.float32)
    ref_exp = np.array([O.exp(float(x)) for x in xs_exp], np.float32)
    ref_sc = np.array([O.sincos_turn(float(x)) for x in xs_sc], np.float32)
    np.testing.assert_array_equal(got_log.view(np.uint32), ref_log.view(np.uint32))
    np.testing.assert_array_equal(got_exp.view(np.uint32), ref_exp.view(np.uint32))
    np.testing.assert_array_equal(got_sc.view(np.uint32), ref_sc.view(np.uint32))


def random_fresnel_cases(n, seed=3):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    flip = np.sum(d * nrm, axis=1) > 0
    nrm[flip] *= -1
    p = rng.normal(size=(n, 3))
    n1 = rng.uniform(1.0, 2.0, n)
    n2 = rng.uniform(1.0, 2.0, n)
    cases = np.concatenate([d, p, nrm, n1[:, None], n2[:, None]], axis=1).astype(np.float32)
    # axis-aligned normals and normal incidence
    cases[:8, 6:9] = [0, 0, -1]
    cases[:4, 0:3] = [0, 0, 1]
    return cases


def test_device_fresnel_bit_exact():
    log("case: device Fresnel vs oracle")
    cases = random_fresnel_cases(3000)
    with host() as h:
        got = h.eval_fresnel(cases)
    ref = O.fresnel(cases)
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("form", ["plain", "persistent"])
def test_synthetic_crystal_bit_exact(crystal, form):
    log(f"case: synthetic crystal {form}")
    src = workloads.synthetic_source(crystal)
    event_ids = np.array([7, 3, 11], np.uint32)
    counts = np.array([40000, 1000, 7], np.uint32)
    ref = O.run_synthetic(crystal.scene, src, 42, event_ids, counts, threads=4)
    with host(42) as h:
        assert h.launch() == LaunchForm()  # a new handle's form
        if form == "persistent":
            h.set_launch(LaunchForm(persistent=True, block_size=256, chunk=256, refill_threshold=16, blocks_per_cu=2))
        else:
            h.set_launch(LaunchForm(persistent=False))
        assert h.launch().persistent == (form == "persistent")
        h.enqueue_synthetic(crystal.scene, src, event_ids, counts)
        out = h.complete_oldest_batch()
    assert out.statistics["dropped_hits"] == 0
    assert_same_hits(out.detections, ref.hits)
    assert_same_stats(out, ref)


@pytest.mark.parametrize("block,refill", [(64, 1), (256, 2), (512, 1), (512, 24)])
def test_ring_generation_passes_bit_exact(crystal, block, refill):
    """The persistent form's ready-photon ring: 8-photon work items of 300
    small events, so that one generation pass draws from many items (photons
    of several events and sources, per lane), refilled from one idle lane on;
    block sizes of 1 to 8 waves give rings of different sizes."""
    log(f"case: ring passes, block {block}, refill {refill}")
    src = workloads.synthetic_source(crystal)
    rng = np.random.default_rng(5)
    event_ids = rng.permutation(1000)[:300].astype(np.uint32)
    counts = rng.integers(0, 60, 300).astype(np.uint32)
    ref = O.run_synthetic(crystal.scene, src, 19, event_ids, counts, threads=4)
    with host(19) as h:
        h.set_launch(LaunchForm(persistent=True, block_size=block, chunk=8, refill_threshold=refill,
                                blocks_per_cu=max(1, 1024 // block)))
        h.enqueue_synthetic(crystal.scene, src, event_ids, counts)
        out = h.complete_oldest_batch()
    assert out.statistics["dropped_hits"] == 0
    assert out.statistics["generated"] == int(counts.sum())
    assert_same_hits(out.detections, ref.hits)
    assert_same_stats(out, ref)


@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("block", [64, 128, 512])
def test_block_sizes_bit_exact(persistent, block):
    """Blocks of 1, 2 and 8 waves on the septa array (a staged BVH and
    property pool ahead of the waves' hit stages in dynamic LDS)."""
    log(f"case: array, block {block}, {'persistent' if persistent else 'plain'}")
    cs = workloads.array_scene()
    em, ev, idx, _ = workloads.array_events(cs, 6, median=3000, seed=3)
    ref = O.run_emissions(cs.scene, em, ev, idx, 61, threads=4)
    with host(61) as h:
        h.set_launch(LaunchForm(persistent=persistent, block_size=block, chunk=256, refill_threshold=16,
                                blocks_per_cu=max(1, 1024 // block)))
        assert h.launch().block_size == block
        h.enqueue_emissions(cs.scene, em, ev, idx)
        out = h.complete_oldest_batch()
    assert out.statistics["dropped_hits"] == 0
    assert_same_hits(out.detections, ref.hits)
    assert_same_stats(out, ref)


def test_emission_crystal_bit_exact(crystal):
    log("case: scintillation emission records")
    em, event_ids, idx = workloads.electron_emissions(crystal, 4, seed=5)
    ref = O.run_emissions(crystal.scene, em, event_ids, idx, 77)
    with host(77) as h:
        h.enqueue_emissions(crystal.scene, em, event_ids, idx)
        out = h.complete_oldest_batch()
    assert_same_hits(out.detections, ref.hits)
    assert_same_stats(out, ref)
    np.testing.assert_array_equal(out.event_statistics["captured"],
                                  np.bincount(idx, weights=em["photon_count"], minlength=4).astype(np.uint64))


def test_array_open_bit_exact():
    """8x8 array without septa: photons leave crystals and cross the world (BVH walk)."""
    log("case: open 8x8 array (BVH)")
    cs = workloads.array_scene(septa=False)
    src = workloads.synthetic_source(cs, crystal_index=27)
    ref = O.run_synthetic(cs.scene, src, 9, [0], [20000], threads=4)
    with host(9) as h:
        h.enqueue_synthetic(cs.scene, src, np.array([0], np.uint32), np.array([20000], np.uint32))
        out = h.complete_oldest_batch()
    assert out.statistics["escaped"] > 0
    assert_same_hits(out.detections, ref.hits)
    assert_same_stats(out, ref)


def test_full_size_invariants(crystal):
    """1e8 photons in one batch: conservation and histogram/hit consistency."""
    log("case: 1e8 synthetic photons, invariants")
    src = workloads.synthetic_source(crystal)
    with host(2024, sort_hits=False) as h:
        h.set_launch(LaunchForm(persistent=True, block_size=256, chunk=1024, refill_threshold=16, blocks_per_cu=2))
        h.enqueue_synthetic(crystal.scene, src, np.array([0], np.uint32), np.array([100_000_000], np.uint32))
        out, ptr = h.complete_device()
    st = out.statistics
    ev = out.event_statistics[0]
    assert st["generated"] == 100_000_000
    total = sum(int(ev[k]) for k in ("detected", "absorbed", "escaped", "truncated", "invalid"))
    assert total == 100_000_000
    assert st["n_hits"] == int(ev["detected"]) == int(out.histogram.sum())
    assert st["dropped_hits"] == 0 and ptr != 0
    frac = int(ev["detected"]) / 1e8
    assert 0.2 < frac < 0.35, frac


def test_photon_trace_invariants(crystal):
    """Per-step device trace: unit direction, polarisation perpendicular to it,
    non-decreasing time, no repeated zero-length steps, and the end point of a
    detected photon equals its record in the batch output."""
    log("case: device photon traces")
    src = workloads.synthetic_source(crystal)
    ref = O.run_synthetic(crystal.scene, src, 42, [3], [64])
    with host(42) as h:
        for pid in range(64):
            t = h.trace_synthetic(crystal.scene, src, 3, pid, 4096)
            assert len(t) > 0
            d, p = t[:, 3:6].astype(np.float64), t[:, 6:9].astype(np.float64)
            assert np.max(np.abs(np.sum(d * d, axis=1) - 1)) < 1e-5
            assert np.max(np.abs(np.sum(d * p, axis=1))) < 1e-5
            assert np.all(np.diff(t[:, 9]) >= 0)
            zero = t[:, 11] > 0
            assert not np.any(zero[1:] & zero[:-1]), f"photon {pid}: consecutive zero-length steps"
            if t[-1, 10] == 1:  # detected
                rec = ref.hits[ref.hits["photon_id"] == pid]
                assert len(rec) == 1
                np.testing.assert_array_equal(t[-1, 0:3], rec["position_mm"][0])
            else:
                assert pid not in set(ref.hits["photon_id"].tolist())


def test_batched_graph_replays_and_per_event_launches(crystal):
    """Uneven events: the persistent form replayed from one captured hipGraph
    over several different batches (with eager work in between) and the plain
    per-event launches give the same results as an eager plain launch."""
    log("case: uneven events, graph replays and per-event launches")
    src = workloads.synthetic_source(crystal)
    batches = [workloads.lognormal_counts(n, 300, sigma=1.5, lo=1, hi=20000, seed=s) for n, s in
               ((40, 1), (25, 2), (40, 3))]
    with host(5) as eager, host(5) as graphed, host(5) as per_event:
        graphed.set_launch(LaunchForm(persistent=True, block_size=256, chunk=256, refill_threshold=16,
                                      blocks_per_cu=4, graph=True))
        per_event.set_launch(LaunchForm(persistent=False, per_event=True))
        for k, counts in enumerate(batches):
            ev = np.arange(100 * k, 100 * k + counts.size, dtype=np.uint32)
            eager.enqueue_synthetic(crystal.scene, src, ev, counts)
            want = eager.complete_oldest_batch()
            for h in (graphed, per_event):
                h.enqueue_synthetic(crystal.scene, src, ev, counts)
                got = h.complete_oldest_batch()
                assert_same_hits(got.detections, want.detections)
                assert got.event_statistics.tobytes() == want.event_statistics.tobytes()
                np.testing.assert_array_equal(got.histogram, want.histogram)
        ref = O.run_synthetic(crystal.scene, src, 5, ev, batches[-1], threads=4)
    assert_same_hits(want.detections, ref.hits)


def forest_scene(n_side=7, seed=4):
    """Air world holding n_side^3 absorbing, scattering glass boxes of random
    sizes on a jittered grid (every fifth box a sensor with a partly reflecting
    skin): photons start below the forest and walk the daughter BVH at every
    world step."""
    from g4go_amd.scene import Material, PropertyTable, Scene, Surface, SurfaceType, Volume

    rng = np.random.default_rng(seed)
    sc = Scene()
    air = sc.add_material(Material(name="air", rindex=PropertyTable.constant_value(1.0)))
    glass = sc.add_material(Material(name="glass", rindex=PropertyTable.constant_value(1.5),
                                     abs_length_mm=PropertyTable.constant_value(30.0),
                                     rayleigh_length_mm=PropertyTable.constant_value(80.0)))
    det = sc.add_surface(Surface(name="det", type=SurfaceType.DielectricMetal,
                                 reflectivity=PropertyTable.constant_value(0.2),
                                 efficiency=PropertyTable.constant_value(0.6)))
    world = sc.add_volume(Volume(name="world", material_id=air, half_mm=(60, 60, 60)))
    sc.world_volume_id = world
    cell = 80.0 / n_side
    k = 0
    for i in range(n_side):
        for j in range(n_side):
            for m in range(n_side):
                half = rng.uniform(0.15, 0.45, 3) * cell
                c = np.array([-40 + (i + 0.5) * cell, -40 + (j + 0.5) * cell, -40 + (m + 0.5) * cell])
                c += rng.uniform(-1, 1, 3) * (0.5 * cell - half)
                sensor = k % 5 == 0
                sc.add_volume(Volume(name=f"b{k}", material_id=glass, parent_volume_id=world,
                                     skin_surface_id=det if sensor else abi.INVALID_ID,
                                     sensor_id=k if sensor else abi.INVALID_ID, center_mm=tuple(c),
                                     half_mm=tuple(half)))
                k += 1
    src = abi.SyntheticSource()
    src.lo_mm[:] = [-45.0, -45.0, -50.0]
    src.hi_mm[:] = [45.0, 45.0, -45.0]
    src.wavelength_mean_nm, src.wavelength_sigma_nm, src.decay_time_ns = 420.0, 20.0, 5.0
    src.volume_id = world
    return sc, src


def test_box_forest_bvh_bit_exact():
    """343 daughters (LDS-staged BVH): hits and counters bit-exact against the oracle."""
    log("case: 343-box forest, BVH walk in LDS")
    sc, src = forest_scene()
    ref = O.run_synthetic(sc, src, 9, [0, 1], [20000, 7000], threads=4)
    with host(9) as h:
        h.set_launch(LaunchForm(persistent=True, block_size=256, chunk=256, refill_threshold=16, blocks_per_cu=4))
        h.enqueue_synthetic(sc, src, np.array([0, 1], np.uint32), np.array([20000, 7000], np.uint32))
        out = h.complete_oldest_batch()
    assert out.statistics["detected"] > 500 and out.statistics["escaped"] > 1000
    assert_same_hits(out.detections, ref.hits)
    assert_same_stats(out, ref)


def test_emission_launch_order_with_shuffled_records(crystal):
    """Records of several events interleaved in the batch: detections come back in
    launch order (record order, then photon order), bit-exact with the oracle."""
    log("case: shuffled emission records, launch order")
    em, event_ids, idx = workloads.electron_emissions(crystal, 6, seed=8)
    perm = np.random.default_rng(2).permutation(em.size)
    em, idx = em[perm].copy(), idx[perm].copy()
    ref = O.run_emissions(crystal.scene, em, event_ids, idx, 19)
    with host(19) as h:
        h.enqueue_emissions(crystal.scene, em, event_ids, idx)
        out = h.complete_oldest_batch()
    assert_same_hits(out.detections, ref.hits)
    assert_same_stats(out, ref)
    # launch order: the record of each hit never decreases along the list
    first = {(int(r["event_id"]), int(r["first_photon_id"])): k for k, r in enumerate(em)}
    rec = []
    for hrec in out.detections:
        cands = [k for (e, f), k in first.items() if e == hrec["event_id"] and f <= hrec["photon_id"]
                 < f + em[k]["photon_count"]]
        rec.append(cands[0])
    assert all(a <= b for a, b in zip(rec, rec[1:]))


def test_device_memory_interop(crystal):
    """Records and indices from torch device tensors, batches ordered after the
    caller's torch stream, hits viewed as a torch tensor: the same detections
    as the host path."""
    import torch

    log("case: device-memory interop (torch tensors, caller stream)")
    em, event_ids, idx = workloads.electron_emissions(crystal, 5, seed=9)
    with host(27) as h:
        h.enqueue_emissions(crystal.scene, em, event_ids, idx)
        want = h.complete_oldest_batch()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        d_em = torch.from_numpy(em.view(np.uint8).copy()).to("cuda", non_blocking=True)
        d_idx = torch.from_numpy(idx.astype(np.int32)).to("cuda", non_blocking=True)
    with host(27) as h:
        h.set_stream(stream)
        h.enqueue_emissions_device(crystal.scene, d_em, event_ids, d_idx)
        out, ptr = h.complete_device()
        hits = h.hits_tensor(ptr, int(out.statistics["n_hits"])).cpu().numpy()
        h.set_stream(None)
    got = np.frombuffer(hits.tobytes(), dtype=abi.HIT_DTYPE)
    assert len(got) == len(want.detections) > 1000
    assert_same_hits(got, want.detections)
    assert out.event_statistics.tobytes() == want.event_statistics.tobytes()
