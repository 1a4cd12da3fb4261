"""Uneven events at full size: 4096 events with log-normal photon counts
(median 5e3, sigma 1.5, clipped to 10..1e6, seed 2024: 6.0e7 photons) in the
wrapped crystal, as the batching benchmark runs them.

* The persistent form replayed from a captured hipGraph, one plain launch
  over the whole batch and 4096 plain per-event launches return the same hit
  list (launch order), per-event counters and histogram.
* Size-independent invariants: every photon ends in exactly one category, and
  hits = detections = histogram total.
* The first 40 events are bit-exact against the FP32 oracle: each photon's
  Philox stream is keyed by (seed, event, photon, step), so a subset of the
  batch replays exactly on its own.
"""
import sys

import numpy as np
import pytest

from g4go_amd import workloads
from g4go_amd.transport import HipTransportHost, LaunchForm, PhotonTransportConfig
from oracle import oracle as O
from tests.test_gpu_parity import assert_same_hits

pytestmark = pytest.mark.gpu

SEED = 2024
N_EVENTS = 4096
N_ORACLE_EVENTS = 40
STAT_FIELDS = ("detected", "absorbed", "escaped", "truncated", "max_bounce", "invalid", "zero_step")


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_uneven_events_full_size(crystal):
    src = workloads.synthetic_source(crystal)
    counts = workloads.lognormal_counts(N_EVENTS, 5000, sigma=1.5, lo=10, hi=1_000_000, seed=SEED)
    ev = np.arange(N_EVENTS, dtype=np.uint32)
    total = int(counts.astype(np.int64).sum())
    forms = {
        "graph": LaunchForm(persistent=True, block_size=256, chunk=1024, refill_threshold=24, blocks_per_cu=6,
                            graph=True),
        "plain_batch": LaunchForm(persistent=False),
        "plain_per_event": LaunchForm(persistent=False, per_event=True),
    }
    outs = {}
    with HipTransportHost(PhotonTransportConfig(seed=SEED, sort_hits=True)) as h:
        for name, form in forms.items():
            log(f"case: uneven events ({N_EVENTS} events, {total} photons), {name}")
            h.set_launch(form)
            h.enqueue_synthetic(crystal.scene, src, ev, counts)
            if name == "graph":  # replay the captured graph a second time
                h.complete_oldest_batch()
                h.enqueue_synthetic(crystal.scene, src, ev, counts)
            outs[name] = h.complete_oldest_batch()
    want = outs["graph"]
    st = want.statistics
    assert st["generated"] == total and st["dropped_hits"] == 0
    ends = sum(int(want.event_statistics[k].sum()) for k in ("detected", "absorbed", "escaped", "truncated",
                                                             "invalid"))
    assert ends == total
    assert len(want.detections) == st["detected"] == int(want.histogram.sum())
    np.testing.assert_array_equal(want.event_statistics["detected"].astype(np.int64),
                                  np.bincount(np.searchsorted(ev, want.detections["event_id"]), minlength=N_EVENTS))
    for name in ("plain_batch", "plain_per_event"):
        log(f"compare {name} with the graph replay")
        got = outs[name]
        assert_same_hits(got.detections, want.detections)
        assert got.event_statistics.tobytes() == want.event_statistics.tobytes()
        np.testing.assert_array_equal(got.histogram, want.histogram)
        assert got.statistics["total_steps"] == st["total_steps"]

    log(f"compare the first {N_ORACLE_EVENTS} events with the FP32 oracle")
    sub = counts[:N_ORACLE_EVENTS]
    ref = O.run_synthetic(crystal.scene, src, SEED, ev[:N_ORACLE_EVENTS], sub, threads=8)
    mine = want.detections[want.detections["event_id"] < N_ORACLE_EVENTS]
    assert_same_hits(mine, ref.hits)
    for k, name in enumerate(STAT_FIELDS):
        np.testing.assert_array_equal(want.event_statistics[name][:N_ORACLE_EVENTS],
                                      ref.stats[:, k].astype(np.uint64), err_msg=name)
