"""The launch autotuner (g4go_amd/autotune.py).

CPU: the candidate space (block sizes of 2, 4 and 8 waves, 8 to 32 waves per
CU) and the cache round trip.  GPU: tune() over three block sizes on a
crystal batch keeps the fastest candidate of its own timing table, caches it,
leaves it set on the host, and every candidate transports the batch to the
same counts (the launch form does not change the physics).
"""
import sys

import numpy as np
import pytest

from g4go_amd import autotune, workloads
from g4go_amd.transport import HipTransportHost, LaunchForm, PhotonTransportConfig


def test_candidate_space():
    forms = list(autotune.candidates())
    assert {f.block_size for f in forms} == {128, 256, 512}
    assert all(512 <= f.block_size * f.blocks_per_cu <= 2048 for f in forms)
    assert all(f.persistent and not f.graph and not f.per_event for f in forms)
    small = {"block_size": (64, 256), "blocks_per_cu": (1, 4, 32), "chunk": (256,), "refill_threshold": (24,)}
    # 64 x 1 and 64 x 4 ask for fewer than 8 waves per CU, 256 x 32 for more than 32
    assert [(f.block_size, f.blocks_per_cu) for f in autotune.candidates(small)] == [(64, 32), (256, 4)]


def test_cache_round_trip(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "CACHE", tmp_path / "cache.json")
    form = LaunchForm(persistent=True, block_size=512, chunk=1024, refill_threshold=16, blocks_per_cu=2)
    table = [{**form.__dict__, "kernel_ms": 2.0}, {**LaunchForm().__dict__, "kernel_ms": 3.0}]
    autotune.store("gfx950/256CU", "crystal", form, table)
    assert autotune.cached("gfx950/256CU", "crystal") == form
    assert autotune.cached("gfx950/256CU", "array") is None
    assert autotune.cached("gfx942/304CU", "crystal") is None


@pytest.mark.gpu
def test_tune_keeps_fastest_candidate(tmp_path, monkeypatch):
    print("case: autotune over block sizes 128, 256, 512 on a 4e6-photon crystal batch", flush=True)
    sys.stdout.flush()
    monkeypatch.setattr(autotune, "CACHE", tmp_path / "cache.json")
    cs = workloads.crystal_scene()
    src = workloads.synthetic_source(cs)
    space = {"block_size": (128, 256, 512), "blocks_per_cu": (4,), "chunk": (256,), "refill_threshold": (24,)}
    counts = []
    with HipTransportHost(PhotonTransportConfig(seed=5, sort_hits=False)) as h:
        h.prepare_scene(cs.scene)

        def run_batch():
            h.enqueue_synthetic(cs.scene, src, np.array([0], np.uint32), np.array([4_000_000], np.uint32))
            out, _ = h.complete_device()
            counts.append((out.statistics["detected"], out.statistics["total_steps"], out.histogram.tobytes()))
            return out.statistics["kernel_ms"]

        key = autotune.device_key(h)
        form, table = autotune.tuned_launch(h, key, "test-crystal", run_batch, space=space, retune=True)
        assert table is not None and len(table) == 3
        assert sorted(r["block_size"] for r in table) == [128, 256, 512]
        best = min(table, key=lambda r: r["kernel_ms"])
        assert form == LaunchForm(**{k: v for k, v in best.items() if k != "kernel_ms"})
        assert h.launch() == form
        assert autotune.cached(key, "test-crystal") == form
        again, none = autotune.tuned_launch(h, key, "test-crystal", run_batch, space=space)
        assert again == form and none is None
        for r in table:
            print(f"  block {r['block_size']}: {r['kernel_ms']:.3f} ms", flush=True)
    assert len(set(counts)) == 1, "launch forms disagree on the batch's counts"
