"""8x8 crystal array with reflective septa (128 boxes: 64 crystals with their
own materials, 64 SiPMs), uneven electron events as emission records, and the
same array with bare crystals (photons cross between crystals).

* bit-exact hits and counters against the FP32 oracle on a small batch;
* the per-crystal absorption gradient (64 x 32 = 2048 parameters) of the
  device replay equals the oracle's replay;
* 1024 events of log-normal size (median 2e4): the persistent launch under
  hipGraph and per-event plain launches give identical counters;
* the 2048-parameter gradient along quadrant directions (16 crystals' nodes
  scaled together) is within 2% of CRN central differences, and is exactly
  zero for crystals no event touched;
* open array: the replay of paths through three or more materials equals the
  oracle's, and the gradient is within 2% of CRN central differences.
"""
import os
import sys

import numpy as np
import pytest

from g4go_amd import abi, workloads
from g4go_amd.transport import HipTransportHost, LaunchForm, PhotonTransportConfig
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PERSISTENT = LaunchForm(persistent=True, block_size=256, chunk=256, refill_threshold=16, blocks_per_cu=4)
THREADS = max(1, min(16, os.cpu_count() or 1))


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def array():
    return workloads.array_scene()


def layout(cs):
    return [(abi.PARAM_ABS_LENGTH, m) for m in cs.crystal_materials]


def test_array_emissions_bit_exact_and_adjoint(array):
    log("case: 8x8 array with septa, 24 events (median 2000 photons), bit-exact + 2048-parameter replay")
    em, ev, idx, crystals = workloads.array_events(array, 24, median=2000, seed=7)
    ref = O.run_emissions(array.scene, em, ev, idx, 91, threads=THREADS)
    w = np.random.default_rng(5).normal(size=abi.HIST_SLOTS)
    with HipTransportHost(PhotonTransportConfig(seed=91, sort_hits=True)) as h:
        h.set_launch(PERSISTENT)
        h.prepare_scene(array.scene)
        assert h.set_parameter_layout(layout(array)) == 64 * 32
        h.enqueue_emissions(array.scene, em, ev, idx)
        out = h.complete_oldest_batch()
        grad, stats = h.adjoint(w)
    assert out.detections.tobytes() == ref.hits.tobytes()
    for k, name in enumerate(("detected", "absorbed", "escaped", "truncated", "max_bounce", "invalid", "zero_step")):
        np.testing.assert_array_equal(out.event_statistics[name], ref.stats[:, k], err_msg=name)
    np.testing.assert_array_equal(out.histogram, ref.hist)
    assert len(set(out.detections["sensor_id"])) > 5
    assert stats["mismatched"] == 0 and stats["rescored"] == 0 and stats["replayed"] == len(ref.hits)
    g_ref, counters = O.adjoint_emissions(array.scene, em, 91, ref.hits, w, layout(array), 64 * 32, threads=THREADS)
    assert int(counters[0]) == len(ref.hits) and int(counters[1]) == 0
    np.testing.assert_allclose(grad, g_ref, rtol=1e-9, atol=1e-9 * np.max(np.abs(g_ref)))
    untouched = np.setdiff1d(np.arange(64), crystals)
    for c in untouched:
        assert not np.any(grad[32 * c:32 * c + 32])


@pytest.fixture(scope="module")
def c5(array):
    return workloads.array_events(array, 1024, median=2e4, seed=2026)


def test_array_batched_equals_per_event(array, c5):
    em, ev, idx, _ = c5
    log(f"case: array 1024 events, {int(em['photon_count'].sum())} photons, graph vs per-event launches")
    outs = {}
    with HipTransportHost(PhotonTransportConfig(seed=17, sort_hits=False)) as h:
        h.prepare_scene(array.scene)
        for name, form in (("graph", LaunchForm(**{**PERSISTENT.__dict__, "graph": True})),
                           ("per_event", LaunchForm(persistent=False, per_event=True))):
            h.set_launch(form)
            for _ in range(2):  # second run replays the graph
                h.enqueue_emissions(array.scene, em, ev, idx)
                out, _ = h.complete_device()
            outs[name] = out
            log(f"  {name}: kernel {out.statistics['kernel_ms']:.2f} ms, "
                f"{out.statistics['generated'] / out.statistics['transport_time_ms'] / 1e6:.3f} G photons/s")
    a, b = outs["graph"], outs["per_event"]
    for name in ("detected", "absorbed", "escaped", "truncated", "max_bounce", "invalid", "zero_step", "captured"):
        np.testing.assert_array_equal(a.event_statistics[name], b.event_statistics[name], err_msg=name)
    np.testing.assert_array_equal(a.histogram, b.histogram)
    assert a.statistics["total_steps"] == b.statistics["total_steps"]
    assert a.statistics["generated"] == int(em["photon_count"].sum())


def test_array_gradient_finite_differences(array, c5):
    """Quadrant directions over the 2048 per-crystal absorption nodes, eps = 0.05."""
    em, ev, idx, crystals = c5
    log("case: array 2048-parameter gradient vs CRN finite differences")
    with HipTransportHost(PhotonTransportConfig(seed=23, sort_hits=False)) as h:
        h.set_launch(PERSISTENT)
        h.prepare_scene(array.scene)
        h.set_parameter_layout(layout(array))
        theta = h.get_parameters()

        def detected(th):
            h.set_parameters(th)
            h.enqueue_emissions(array.scene, em, ev, idx)
            out, _ = h.complete_device()
            return out

        base = detected(theta)
        fwd_ms = base.statistics["kernel_ms"]
        grad, stats = h.adjoint(np.ones(abi.HIST_SLOTS))
        assert stats["mismatched"] == 0 and stats["rescored"] == 0
        log(f"  forward {fwd_ms:.2f} ms, adjoint {stats['kernel_ms']:.2f} ms ({stats['replayed']} replays)")
        for c in np.setdiff1d(np.arange(64), crystals):
            assert not np.any(grad[32 * c:32 * c + 32])
        i = np.arange(64) // 8
        j = np.arange(64) % 8
        for q in range(4):
            members = np.nonzero(((i >= 4).astype(int) * 2 + (j >= 4).astype(int)) == q)[0]
            v = np.zeros_like(theta)
            for c in members:
                v[32 * c:32 * c + 32] = theta[32 * c:32 * c + 32]
            n_p = detected(theta + 0.05 * v).statistics["detected"]
            n_m = detected(theta - 0.05 * v).statistics["detected"]
            fd = (n_p - n_m) / 0.1
            ad = float(grad @ v)
            log(f"  quadrant {q}: adjoint {ad:.6g} finite difference {fd:.6g} ratio {ad / fd:.5f}")
            assert abs(ad / fd - 1) < 0.02, q
        h.set_parameters(theta)


@pytest.fixture(scope="module")
def open_array():
    return workloads.array_scene(septa=False)


def test_open_array_replay_of_long_paths(open_array):
    """Bare crystals: photons leave a crystal into the air and enter others, so
    many paths cross three or more materials and are scored on the replay's
    second pass; the gradient still equals the oracle's replay."""
    log("case: open 8x8 array, 24 events (median 2000 photons), bit-exact + replay of multi-crystal paths")
    em, ev, idx, _ = workloads.array_events(open_array, 24, median=2000, seed=7)
    ref = O.run_emissions(open_array.scene, em, ev, idx, 93, threads=THREADS)
    w = np.random.default_rng(6).normal(size=abi.HIST_SLOTS)
    with HipTransportHost(PhotonTransportConfig(seed=93, sort_hits=True)) as h:
        h.set_launch(PERSISTENT)
        h.prepare_scene(open_array.scene)
        h.set_parameter_layout(layout(open_array))
        h.enqueue_emissions(open_array.scene, em, ev, idx)
        out = h.complete_oldest_batch()
        grad, stats = h.adjoint(w)
    assert out.detections.tobytes() == ref.hits.tobytes()
    g_ref, counters = O.adjoint_emissions(open_array.scene, em, 93, ref.hits, w, layout(open_array), 64 * 32,
                                          threads=THREADS)
    log(f"  {stats['replayed']} replays, {stats['rescored']} through more than two materials")
    assert stats["mismatched"] == 0 and stats["replayed"] == len(ref.hits) == int(counters[0])
    assert stats["rescored"] == int(counters[2]) > 0.01 * len(ref.hits)
    # the second passes replay some paths twice
    assert int(counters[3]) < stats["steps"] < 2 * int(counters[3])
    np.testing.assert_allclose(grad, g_ref, rtol=1e-9, atol=1e-9 * np.max(np.abs(g_ref)))


def test_open_array_gradient_finite_differences(open_array):
    """The open array's 2048-parameter gradient along quadrant directions
    within 2% of CRN central differences (eps = 0.05)."""
    em, ev, idx, _ = workloads.array_events(open_array, 1024, median=2e4, seed=2026)
    log(f"case: open array 1024 events, {int(em['photon_count'].sum())} photons, gradient vs CRN differences")
    with HipTransportHost(PhotonTransportConfig(seed=29, sort_hits=False)) as h:
        h.set_launch(PERSISTENT)
        h.prepare_scene(open_array.scene)
        h.set_parameter_layout(layout(open_array))
        theta = h.get_parameters()

        def detected(th):
            h.set_parameters(th)
            h.enqueue_emissions(open_array.scene, em, ev, idx)
            out, _ = h.complete_device()
            return out

        base = detected(theta)
        grad, stats = h.adjoint(np.ones(abi.HIST_SLOTS))
        assert stats["mismatched"] == 0 and stats["replayed"] == base.statistics["detected"]
        log(f"  forward {base.statistics['kernel_ms']:.2f} ms, adjoint {stats['kernel_ms']:.2f} ms "
            f"({stats['replayed']} replays, {stats['rescored']} rescored)")
        assert stats["rescored"] > 0
        i = np.arange(64) // 8
        j = np.arange(64) % 8
        for q in range(4):
            members = np.nonzero(((i >= 4).astype(int) * 2 + (j >= 4).astype(int)) == q)[0]
            v = np.zeros_like(theta)
            for c in members:
                v[32 * c:32 * c + 32] = theta[32 * c:32 * c + 32]
            n_p = detected(theta + 0.05 * v).statistics["detected"]
            n_m = detected(theta - 0.05 * v).statistics["detected"]
            fd = (n_p - n_m) / 0.1
            ad = float(grad @ v)
            log(f"  quadrant {q}: adjoint {ad:.6g} finite difference {fd:.6g} ratio {ad / fd:.5f}")
            assert abs(ad / fd - 1) < 0.02, q
        h.set_parameters(theta)
