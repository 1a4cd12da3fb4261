"""Batch slots: up to max_in_flight batches enqueued before the oldest is
completed, completed in enqueue order (CPU checks use photon-less batches, which
never reach the device; the device check compares pipelined with one-at-a-time
batches bit for bit)."""
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
def test_pipelined_batches_equal_sequential(crystal):
    """Three slots with up to three batches in flight give the same hits and
    counters as one slot, batch by batch."""
    print("case: pipelined emission batches (3 slots) vs one slot", flush=True)
    em, ev, idx = workloads.electron_emissions(crystal, 12, seed=6)
    batches = []
    for b in range(4):
        sel = (idx >= 3 * b) & (idx < 3 * b + 3)
        batches.append((em[sel], ev[3 * b:3 * b + 3], (idx[sel] - 3 * b).astype(np.uint32)))
    form = LaunchForm(persistent=True, block_size=256, chunk=256, refill_threshold=16, blocks_per_cu=4, graph=True)
    seq = []
    with HipTransportHost(PhotonTransportConfig(seed=21)) as h:
        h.set_launch(form)
        for b in batches:
            h.enqueue_emissions(crystal.scene, *b)
            seq.append(h.complete_oldest_batch())
    pipe = []
    with HipTransportHost(PhotonTransportConfig(seed=21, max_in_flight=3)) as h:
        h.set_launch(form)
        h.prepare_scene(crystal.scene)
        for b in batches[:3]:
            h.enqueue_emissions(crystal.scene, *b)
        pipe.append(h.complete_oldest_batch())
        h.enqueue_emissions(crystal.scene, *batches[3])
        pipe += [h.complete_oldest_batch() for _ in range(3)]
    for a, b in zip(seq, pipe):
        assert a.detections.tobytes() == b.detections.tobytes()
        assert a.event_statistics.tobytes() == b.event_statistics.tobytes()
        np.testing.assert_array_equal(a.histogram, b.histogram)
