"""Autotuner of the persistent transport launch.

Sweeps the launch form of the persistent kernel -- block size, resident
blocks per CU (grid depth), photons per queue item and the lane-refill
threshold -- on a sample of the caller's workload, timing each candidate by
the transport kernel's own HIP-event duration, and keeps the fastest.  Results
are cached per device and workload tag in autotune_cache.json next to this
file.  The device key is the architecture name and compute-unit count
(device_key), so that boxes reporting different marketing names share it.
"""
from __future__ import annotations

import itertools
import json
import time
from pathlib import Path
from typing import Callable, Iterable, Optional

from .transport import HipTransportHost, LaunchForm

CACHE = Path(__file__).resolve().parent / "autotune_cache.json"

# Block sizes of 2, 4 and 8 waves: each wave's hit stage is sized with the
# block in dynamic LDS, so the block size trades the per-block scene copy
# (staged once per block) against the granularity of residency on a CU.
# Candidates ask for 8 to 32 waves per CU (the kernels run at most 16,
# four per SIMD, so deeper grids only queue blocks behind the resident ones).
DEFAULT_SPACE = {
    "block_size": (128, 256, 512),
    "blocks_per_cu": (2, 4, 6, 8, 12, 16),
    "chunk": (256, 1024, 4096),
    "refill_threshold": (2, 4, 8, 16, 24),
}


def candidates(space: Optional[dict] = None) -> Iterable[LaunchForm]:
    space = space or DEFAULT_SPACE
    for bs, bpc, chunk, refill in itertools.product(space["block_size"], space["blocks_per_cu"], space["chunk"],
                                                    space["refill_threshold"]):
        if not 512 <= bs * bpc <= 2048:  # 8 to 32 waves per CU
            continue
        yield LaunchForm(persistent=True, block_size=bs, chunk=chunk, refill_threshold=refill, blocks_per_cu=bpc)


def tune(host: HipTransportHost, run_batch: Callable[[], float], space: Optional[dict] = None,
         repeats: int = 2) -> tuple[LaunchForm, list[dict]]:
    """Times run_batch() (which enqueues and completes one batch and returns the
    kernel milliseconds) for every candidate; returns the best form and the table."""
    table = []
    best, best_ms = None, float("inf")
    for form in candidates(space):
        host.set_launch(form)
        run_batch()  # warm-up
        ms = min(run_batch() for _ in range(repeats))
        table.append({**form.__dict__, "kernel_ms": ms})
        if ms < best_ms:
            best, best_ms = form, ms
    host.set_launch(best)
    return best, table


def device_key(host: HipTransportHost) -> str:
    """'<gcnArchName>/<CUs>CU' of the host's device, e.g. 'gfx950:sramecc+:xnack-/256CU'."""
    name, cu = host.device_info()
    return f"{name.split()[0] if name else 'unknown'}/{cu}CU"


def _load_cache() -> dict:
    if CACHE.is_file():
        try:
            return json.loads(CACHE.read_text())
        except (OSError, ValueError):
            return {}
    return {}


def cached(device: str, tag: str) -> Optional[LaunchForm]:
    entry = _load_cache().get(f"{device}|{tag}")
    if entry is None:
        return None
    return LaunchForm(**entry["launch"])


def store(device: str, tag: str, form: LaunchForm, table: list[dict]) -> None:
    data = _load_cache()
    best = min(table, key=lambda r: r["kernel_ms"])
    data[f"{device}|{tag}"] = {"launch": form.__dict__, "kernel_ms": best["kernel_ms"], "time": time.time(),
                               "candidates": len(table)}
    CACHE.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")


def tuned_launch(host: HipTransportHost, device: str, tag: str, run_batch: Callable[[], float],
                 space: Optional[dict] = None, retune: bool = False) -> tuple[LaunchForm, Optional[list]]:
    """Cached form for (device, tag), tuning and caching it when absent.
    Returns the form and the timing table (None when it came from the cache)."""
    form = None if retune else cached(device, tag)
    table = None
    if form is None:
        form, table = tune(host, run_batch, space)
        store(device, tag, form, table)
    host.set_launch(form)
    return form, table
