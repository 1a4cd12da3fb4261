"""Statistical gates of the device (FP32 state, FP64 time) against the FP64
oracle on 1 MeV electron events in the wrapped crystal (30 photons/keV,
scintillation emission records).

Same Philox streams (seed 77, 500 events): the FP64 oracle replays every
photon with the device's draws, so the two differ only where FP32 rounding
changes a decision.  Per-event detected mean and width agree within 0.5%
(the accuracy gate of the path), and the arrival-time distributions pass a two-sample KS
test at p > 0.01.

Independent streams (device seed 78 on all 5000 events, oracle seed 77 on
500): the mean per-event count agrees within 0.5% (its standard error is
about 0.06%), the width within three standard errors of the difference of
two sample widths (about 4.5% at 500 events: a 0.5% width check would need
~1e5 oracle events), and the arrival times pass KS at p > 0.01; the
reference's own CPU/GPU gates are 5% mean, 10% width, KS p >= 0.01
(test/TestOpticalTransport.cxx:320-336).
"""
import os

import numpy as np
import pytest
from scipy import stats as ss

from g4go_amd import workloads
from g4go_amd.transport import HipTransportHost, LaunchForm, PhotonTransportConfig

pytestmark = pytest.mark.gpu

ORACLE_EVENTS = 500
ALL_EVENTS = 5000
PERSISTENT = LaunchForm(persistent=True, block_size=256, chunk=256, refill_threshold=16, blocks_per_cu=4)


def log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def c1():
    cs = workloads.crystal_scene()
    em, ev, idx = workloads.electron_emissions(cs, ALL_EVENTS, seed=2025)
    return cs, em, ev, idx


@pytest.fixture(scope="module")
def oracle64(c1):
    from oracle import oracle as O

    cs, em, ev, idx = c1
    sel = idx < ORACLE_EVENTS
    log(f"oracle FP64: {ORACLE_EVENTS} events, {int(em['photon_count'][sel].sum())} photons")
    threads = max(1, min(16, os.cpu_count() or 1))
    return O.run_emissions(cs.scene, em[sel], ev[:ORACLE_EVENTS], idx[sel], 77, precision=64, threads=threads)


def device_run(c1, seed, n_events):
    cs, em, ev, idx = c1
    sel = idx < n_events
    with HipTransportHost(PhotonTransportConfig(seed=seed, sort_hits=True)) as h:
        h.set_launch(PERSISTENT)
        h.enqueue_emissions(cs.scene, em[sel], ev[:n_events], idx[sel])
        return h.complete_oldest_batch()


def moments(x):
    x = np.asarray(x, np.float64)
    return x.mean(), x.std(ddof=1)


def test_same_streams_fp32_vs_fp64(c1, oracle64):
    log("case: C1 same streams, device FP32 vs oracle FP64")
    out = device_run(c1, 77, ORACLE_EVENTS)
    dev = out.event_statistics["detected"]
    ref = oracle64.stats[:, 0]
    (m_d, s_d), (m_r, s_r) = moments(dev), moments(ref)
    changed = int(np.abs(dev.astype(np.int64) - ref.astype(np.int64)).sum())
    log(f"  mean {m_d:.3f} vs {m_r:.3f}, width {s_d:.3f} vs {s_r:.3f}, |per-event differences| summed {changed}")
    assert abs(m_d / m_r - 1) < 0.005
    assert abs(s_d / s_r - 1) < 0.005
    ks = ss.ks_2samp(out.detections["time_ns"], oracle64.hits["time_ns"])
    log(f"  TOF KS statistic {ks.statistic:.2e}, p {ks.pvalue:.3f}")
    assert ks.pvalue > 0.01


def test_independent_streams(c1, oracle64):
    log("case: C1 independent streams, device 5000 events vs oracle FP64 500 events")
    out = device_run(c1, 78, ALL_EVENTS)
    dev = out.event_statistics["detected"]
    ref = oracle64.stats[:, 0]
    (m_d, s_d), (m_r, s_r) = moments(dev), moments(ref)
    se_mean = np.sqrt(s_d**2 / dev.size + s_r**2 / ref.size)
    se_width = np.sqrt(s_d**2 / (2 * (dev.size - 1)) + s_r**2 / (2 * (ref.size - 1)))
    log(f"  mean {m_d:.2f} vs {m_r:.2f} (se {se_mean:.2f}), width {s_d:.2f} vs {s_r:.2f} (se {se_width:.2f})")
    assert abs(m_d / m_r - 1) < 0.005
    assert abs(s_d - s_r) < 3 * se_width
    ks = ss.ks_2samp(out.detections["time_ns"], oracle64.hits["time_ns"])
    log(f"  TOF KS statistic {ks.statistic:.2e}, p {ks.pvalue:.3f} ({len(out.detections)} vs {len(oracle64.hits)})")
    assert ks.pvalue > 0.01
    ks_n = ss.ks_2samp(dev, ref)
    log(f"  per-event count KS p {ks_n.pvalue:.3f}")
    assert ks_n.pvalue > 0.01
