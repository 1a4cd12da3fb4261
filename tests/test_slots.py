"""Batch slots: up to max_in_flight batches enqueued before the oldest is
completed, completed in enqueue order (CPU checks use photon-less batches, which
never reach the device; the device check compares every pipelined batch with
the oracle's run of that batch, bit for bit)."""
import numpy as np
import pytest

from g4go_amd import abi, workloads
from g4go_amd.transport import HipTransportHost, LaunchForm, PhotonTransportConfig


def empty_batch(cs, event_id):
    em = np.zeros(1, abi.EMISSION_DTYPE)
    em["event_id"] = event_id
    em["volume_id"] = cs.crystal_ids[0]
    em["photon_count"] = 0
    return em, np.array([event_id], np.uint32), np.zeros(1, np.uint32)


def test_ring_order_and_limits(crystal):
    with HipTransportHost(PhotonTransportConfig(max_in_flight=3)) as h:
        h.prepare_scene(crystal.scene)
        for e in (11, 12, 13):
            h.enqueue_emissions(crystal.scene, *empty_batch(crystal, e))
        with pytest.raises(Exception, match="queue is full"):
            h.enqueue_emissions(crystal.scene, *empty_batch(crystal, 14))
        with pytest.raises(Exception, match="pending batch"):
            h.set_launch(LaunchForm(persistent=True))
        got = [int(h.complete_oldest_batch().event_statistics["event_id"][0]) for _ in range(2)]
        h.enqueue_emissions(crystal.scene, *empty_batch(crystal, 15))
        got += [int(h.complete_oldest_batch().event_statistics["event_id"][0]) for _ in range(2)]
        assert got == [11, 12, 13, 15]
        with pytest.raises(Exception, match="queue is empty"):
            h.complete_oldest_batch()


def test_single_slot_is_the_reference_behaviour(crystal):
    with HipTransportHost(PhotonTransportConfig()) as h:
        h.prepare_scene(crystal.scene)
        h.enqueue_emissions(crystal.scene, *empty_batch(crystal, 1))
        with pytest.raises(Exception, match="queue is full"):
            h.enqueue_emissions(crystal.scene, *empty_batch(crystal, 2))


def test_in_flight_limit():
    with pytest.raises(Exception, match="at most 4"):
        HipTransportHost(PhotonTransportConfig(max_in_flight=5))


@pytest.mark.gpu
@pytest.mark.parametrize("slots", [2, 3, 4])
def test_pipelined_batches_match_oracle(crystal, slots):
    """Up to `slots` batches of different sizes in flight (scintillation
    records, and one synthetic batch between them): every batch, completed in
    enqueue order, equals the FP32 oracle's run of that batch alone, bit for
    bit (hits in launch order, per-event counters, histogram)."""
    from oracle import oracle as O
    from tests.test_gpu_parity import assert_same_hits, assert_same_stats

    print(f"case: pipelined batches, {slots} slots, vs the oracle", flush=True)
    em, ev, idx = workloads.electron_emissions(crystal, 16, seed=6)
    src = workloads.synthetic_source(crystal)
    batches, refs = [], []
    first = 0
    for n in (1, 4, 2, None, 5, 1, 3):
        if n is None:  # a synthetic batch between the record batches
            sev, scounts = np.array([90, 91], np.uint32), np.array([20000, 700], np.uint32)
            batches.append(("synthetic", (crystal.scene, src, sev, scounts)))
            refs.append(O.run_synthetic(crystal.scene, src, 21, sev, scounts, threads=4))
            continue
        sel = (idx >= first) & (idx < first + n)
        b = (em[sel], ev[first:first + n], (idx[sel] - first).astype(np.uint32))
        batches.append(("records", (crystal.scene, *b)))
        refs.append(O.run_emissions(crystal.scene, *b, 21, threads=4))
        first += n
    form = LaunchForm(persistent=True, block_size=256, chunk=256, refill_threshold=16, blocks_per_cu=4, graph=True)
    outs = []
    with HipTransportHost(PhotonTransportConfig(seed=21, max_in_flight=slots)) as h:
        h.set_launch(form)
        h.prepare_scene(crystal.scene)
        pending = 0
        for kind, args in batches:
            if pending == slots:
                outs.append(h.complete_oldest_batch())
                pending -= 1
            if kind == "records":
                h.enqueue_emissions(*args)
            else:
                h.enqueue_synthetic(*args)
            pending += 1
        while pending:
            outs.append(h.complete_oldest_batch())
            pending -= 1
    assert len(outs) == len(refs)
    for k, (out, ref) in enumerate(zip(outs, refs)):
        print(f"  batch {k}: {len(out.detections)} hits", flush=True)
        assert_same_hits(out.detections, ref.hits)
        assert_same_stats(out, ref)
