"""HipTransportHost: the reference backend interface over the C ABI.

Mirrors G4GO::Optical::OptiXTransportHost
(include/g4go/optical/optix/OptiXTransportHost.hpp:10-26): a constructor
taking the transport configuration, PrepareScene, EnqueueEmissions and
CompleteOldestBatch, with one in-flight batch.  C-ABI status codes are raised
as the reference's exception kinds: runtime_error -> RuntimeError,
invalid_argument -> ValueError, overflow_error -> OverflowError,
logic_error -> LogicError.
"""
from __future__ import annotations

import ctypes as C
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import abi
from .scene import Scene


# Detections are copied out of the library's pinned buffer in parallel slices
# when the batch is large: ctypes releases the GIL during memmove, and one core
# copies ~25 GB/s, slower than the device-to-host transfer itself.
_COPY_SLICE_BYTES = 4 << 20
_COPY_THREADS = 4
_copy_pool: Optional[ThreadPoolExecutor] = None


def _copy_out(dst: int, src: int, nbytes: int) -> None:
    global _copy_pool
    if nbytes < 2 * _COPY_SLICE_BYTES:
        C.memmove(dst, src, nbytes)
        return
    if _copy_pool is None:
        _copy_pool = ThreadPoolExecutor(_COPY_THREADS, thread_name_prefix="g4go-copy")
    step = -(-nbytes // _COPY_THREADS)
    futures = [_copy_pool.submit(C.memmove, dst + o, src + o, min(step, nbytes - o)) for o in range(0, nbytes, step)]
    for f in futures:
        f.result()


class LogicError(RuntimeError):
    """std::logic_error of the reference (queue full or empty, no scene)."""


_ERRORS = {
    abi.G4GO_ERROR_RUNTIME: RuntimeError,
    abi.G4GO_ERROR_INVALID_ARGUMENT: ValueError,
    abi.G4GO_ERROR_OVERFLOW: OverflowError,
    abi.G4GO_ERROR_LOGIC: LogicError,
}


@dataclass
class PhotonTransportConfig:
    """PhotonTransportConfig (include/g4go/optical/PhotonTransportConfig.hpp:26-40)."""

    seed: int = 0
    max_photons_per_event: int = 5_000_000
    max_bounces_per_photon: int = 4096
    boundary_tolerance_mm: float = 1.0e-4
    target_photons_per_batch: int = 1_000_000
    max_events_per_batch: int = 10_000
    enable_performance_diagnostics: bool = False
    device: int = -1
    sort_hits: bool = True          # detections in launch order
    max_in_flight: int = 1          # batches enqueued before the oldest completes (1..4)

    def to_abi(self) -> abi.Config:
        return abi.Config(self.seed, self.max_photons_per_event, self.max_bounces_per_photon,
                          self.boundary_tolerance_mm, self.target_photons_per_batch, self.max_events_per_batch,
                          int(self.enable_performance_diagnostics), self.device, int(self.sort_hits),
                          int(self.max_in_flight))


@dataclass
class LaunchForm:
    """Launch form of a handle (g4go_launch).  The defaults are those of a new
    handle: the persistent kernel with 512-thread blocks, 256-photon work
    items, refill at 4 idle lanes, 2 blocks per CU, photons generated at
    refill (pregenerate off).  The plain form
    (`LaunchForm(persistent=False)`) runs one wave per item of at most 64
    photons and ignores chunk and refill_threshold."""

    persistent: bool = True
    block_size: int = 512
    chunk: int = 256
    refill_threshold: int = 4
    blocks_per_cu: int = 2
    per_event: bool = False
    graph: bool = False
    pregenerate: bool = False

    def to_abi(self) -> abi.Launch:
        return abi.Launch(int(self.persistent), self.block_size, self.chunk, self.refill_threshold,
                          self.blocks_per_cu, int(self.per_event), int(self.graph), int(self.pregenerate))


@dataclass
class PhotonTransportOutput:
    """PhotonTransportOutput (PhotonTransport.hpp:152-157) plus the arrival-time histogram."""

    detections: np.ndarray                 # abi.HIT_DTYPE records
    statistics: dict                       # batch totals
    event_statistics: np.ndarray           # abi.EVENT_DTYPE records
    histogram: np.ndarray                  # 258 u64: 256 bins over [0, 200) ns, underflow, overflow
    performance: dict = field(default_factory=dict)


class HipTransportHost:
    def __init__(self, configuration: Optional[PhotonTransportConfig] = None) -> None:
        self._lib = abi.load()
        self.configuration = configuration or PhotonTransportConfig()
        self._cfg = self.configuration.to_abi()
        handle = C.c_void_p()
        status = self._lib.g4go_create(C.byref(self._cfg), C.byref(handle))
        if status != abi.G4GO_OK:
            raise _ERRORS.get(status, RuntimeError)(self._lib.g4go_last_error(None).decode())
        self._h = handle
        self.batches_completed = 0  # batches completed so far (a serial for callers such as autograd)
        self._scene: Optional[Scene] = None
        self._scene_keep: list = []
        self._n_params = 0

    # ---- lifecycle ----
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.g4go_destroy(self._h)
            self._h = None

    def __del__(self) -> None:
        self.close()

    def __enter__(self) -> "HipTransportHost":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def _check(self, status: int) -> None:
        if status != abi.G4GO_OK:
            message = self._lib.g4go_last_error(self._h).decode()
            raise _ERRORS.get(status, RuntimeError)(message)

    def set_launch(self, launch: LaunchForm) -> None:
        form = launch.to_abi()
        self._check(self._lib.g4go_set_launch(self._h, C.byref(form)))

    def launch(self) -> LaunchForm:
        form = abi.Launch()
        self._check(self._lib.g4go_get_launch(self._h, C.byref(form)))
        return LaunchForm(bool(form.persistent), form.block_size, form.chunk, form.refill_threshold,
                          form.blocks_per_cu, bool(form.per_event), bool(form.graph), bool(form.pregenerate))

    # ---- the reference interface ----
    def prepare_scene(self, scene: Scene) -> None:
        """PrepareScene: rebuilds when a different Scene object is given (OptiXTransportHost.cpp:511-519)."""
        if scene is self._scene:
            return
        desc, keep = scene.to_abi()
        self._check(self._lib.g4go_prepare_scene(self._h, C.byref(desc)))
        self._scene = scene
        self._scene_keep = keep

    def enqueue_emissions(self, scene: Scene, emissions: np.ndarray, event_ids: np.ndarray,
                          emission_event_indices: np.ndarray) -> None:
        """EnqueueEmissions(scene, {emissions, eventIDs, emissionEventIndices})."""
        self.prepare_scene(scene)
        em = np.ascontiguousarray(emissions, dtype=abi.EMISSION_DTYPE)
        ev = np.ascontiguousarray(event_ids, dtype=np.uint32)
        idx = np.ascontiguousarray(emission_event_indices, dtype=np.uint32)
        self._check(self._lib.g4go_enqueue_emissions(self._h, em.ctypes.data if em.size else None, em.size,
                                                     ev.ctypes.data if ev.size else None, ev.size,
                                                     idx.ctypes.data if idx.size else None))

    # ---- device-memory interop ----
    def set_stream(self, stream) -> None:
        """Orders every later batch after the work queued on `stream` (a
        torch.cuda.Stream, a raw hipStream_t as int, or None to stop)."""
        handle = getattr(stream, "cuda_stream", stream)
        self._check(self._lib.g4go_set_stream(self._h, C.c_void_p(int(handle)) if handle else None))

    def enqueue_emissions_device(self, scene: Scene, emissions, event_ids: np.ndarray, emission_event_indices) -> None:
        """EnqueueEmissions with the records (a device tensor of n x 128 bytes,
        abi.EMISSION_DTYPE layout) and the emission event indices (a device
        tensor of n uint32/int32) in device memory; event IDs on the host."""
        if emissions.numel() and (not emissions.is_cuda or not emission_event_indices.is_cuda):
            raise ValueError("records and indices must be device tensors")
        if not emissions.is_contiguous() or not emission_event_indices.is_contiguous():
            raise ValueError("device tensors must be contiguous")
        nbytes = emissions.numel() * emissions.element_size()
        if nbytes % abi.EMISSION_DTYPE.itemsize:
            raise ValueError("the records tensor must hold whole 128-byte records")
        n = nbytes // abi.EMISSION_DTYPE.itemsize
        if emission_event_indices.numel() != n or emission_event_indices.element_size() != 4:
            raise ValueError("one 4-byte event index per record")
        self.prepare_scene(scene)
        ev = np.ascontiguousarray(event_ids, dtype=np.uint32)
        self._check(self._lib.g4go_enqueue_emissions_device(
            self._h, emissions.data_ptr() if n else None, n, ev.ctypes.data if ev.size else None, ev.size,
            emission_event_indices.data_ptr() if n else None))

    @staticmethod
    def hits_tensor(hits_device: int, n_hits: int):
        """A torch view (uint8, n_hits x 48, abi.HIT_DTYPE records) of the hits
        complete_device left in HBM; valid until the next enqueue."""
        import torch

        class _View:
            __cuda_array_interface__ = {"shape": (n_hits, abi.HIT_DTYPE.itemsize), "typestr": "|u1",
                                        "data": (hits_device, False), "version": 2, "strides": None}

        if n_hits == 0:
            return torch.empty((0, abi.HIT_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
        return torch.as_tensor(_View(), device="cuda")

    def enqueue_synthetic(self, scene: Scene, source: abi.SyntheticSource, event_ids: np.ndarray,
                          photon_counts: np.ndarray) -> None:
        self.prepare_scene(scene)
        ev = np.ascontiguousarray(event_ids, dtype=np.uint32)
        counts = np.ascontiguousarray(photon_counts, dtype=np.uint32)
        if ev.shape != counts.shape:
            raise ValueError("event_ids and photon_counts must have the same length")
        self._check(self._lib.g4go_enqueue_synthetic(self._h, C.byref(source), ev.ctypes.data if ev.size else None,
                                                     counts.ctypes.data if counts.size else None, ev.size))

    def transport_synthetic_noalloc(self, source: abi.SyntheticSource, event_ids, photon_counts,
                                    outputs: abi.DeviceOutputs, stream=None) -> None:
        """g4go_transport_synthetic_noalloc: a synthetic batch that allocates nothing.  event_ids and
        photon_counts are device tensors of n_events 4-byte values; `outputs` holds device pointers
        (abi.DeviceOutputs).  The work is queued on `stream` (a torch.cuda.Stream or raw handle; None:
        the handle's stream) and complete when the stream reaches it."""
        n = int(event_ids.numel())
        if int(photon_counts.numel()) != n:
            raise ValueError("event_ids and photon_counts must have the same length")
        for t in (event_ids, photon_counts):
            if n and (not t.is_cuda or t.element_size() != 4 or not t.is_contiguous()):
                raise ValueError("event_ids and photon_counts must be contiguous 4-byte device tensors")
        handle = getattr(stream, "cuda_stream", stream)
        self._check(self._lib.g4go_transport_synthetic_noalloc(
            self._h, C.byref(source), event_ids.data_ptr() if n else None, photon_counts.data_ptr() if n else None,
            n, C.byref(outputs), C.c_void_p(int(handle)) if handle else None))

    def complete_oldest_batch(self) -> PhotonTransportOutput:
        res = abi.BatchResult()
        self._check(self._lib.g4go_complete(self._h, C.byref(res)))
        self.batches_completed += 1
        return self._output(res, copy_hits=True)

    def complete_device(self) -> tuple[PhotonTransportOutput, int]:
        """Completes the batch leaving detections in HBM; returns (output, device pointer of the hits)."""
        res = abi.BatchResult()
        ptr = C.c_void_p()
        self._check(self._lib.g4go_complete_device(self._h, C.byref(res), C.byref(ptr)))
        self.batches_completed += 1
        return self._output(res, copy_hits=False), int(ptr.value or 0)

    def _output(self, res: abi.BatchResult, copy_hits: bool) -> PhotonTransportOutput:
        n_hits = int(res.n_hits)
        hits = np.empty(n_hits if copy_hits else 0, dtype=abi.HIT_DTYPE)
        if copy_hits and n_hits:
            _copy_out(hits.ctypes.data, C.cast(res.hits, C.c_void_p).value, n_hits * abi.HIT_DTYPE.itemsize)
        events = np.empty(int(res.n_events), dtype=abi.EVENT_DTYPE)
        if res.n_events:
            C.memmove(events.ctypes.data, res.events, int(res.n_events) * abi.EVENT_DTYPE.itemsize)
        hist = np.ctypeslib.as_array(res.histogram, shape=(abi.HIST_SLOTS,)).copy()
        st = res.stats
        stats = {name: getattr(st, name) for name, _ in abi.BatchStats._fields_}
        stats["n_hits"] = n_hits
        perf = {k: stats[k] for k in ("kernel_ms", "host_to_device_ms", "device_to_host_ms", "host_to_device_bytes",
                                      "device_to_host_bytes", "total_steps", "device_memset_ms", "device_order_ms",
                                      "metadata_to_host_ms", "metadata_to_host_bytes", "hits_to_host_ms",
                                      "hits_to_host_bytes")}
        return PhotonTransportOutput(hits, stats, events, hist, perf)

    def device_info(self) -> tuple[str, int]:
        buf = C.create_string_buffer(256)
        cu = C.c_int32()
        self._check(self._lib.g4go_device_info(self._h, buf, 256, C.byref(cu)))
        return buf.value.decode(), cu.value

    def trace_synthetic(self, scene: Scene, source: abi.SyntheticSource, event_id: int, photon_id: int,
                        max_steps: int = 4096) -> np.ndarray:
        """Per-step records (n, 16) of one synthetic photon transported on the device."""
        self.prepare_scene(scene)
        out = np.zeros((max_steps, 16), np.float32)
        n = C.c_uint32()
        self._check(self._lib.g4go_trace_synthetic(self._h, C.byref(source), event_id, photon_id, max_steps,
                                                   out.ctypes.data, C.byref(n)))
        return out[:n.value]

    # ---- differentiation ----
    def set_parameter_layout(self, tables) -> int:
        """Binds (kind, index) property tables to theta (kinds: abi.PARAM_*); returns len(theta)."""
        arr = (abi.ParamTable * max(1, len(tables)))(*[abi.ParamTable(int(k), int(i)) for k, i in tables])
        n = C.c_uint32()
        self._check(self._lib.g4go_set_parameter_layout(self._h, arr, len(tables), C.byref(n)))
        self._n_params = n.value
        return n.value

    def get_parameters(self) -> np.ndarray:
        theta = np.zeros(self._n_params, np.float64)
        self._check(self._lib.g4go_get_parameters(self._h, theta.ctypes.data, theta.size))
        return theta

    def set_parameters(self, theta) -> None:
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        self._check(self._lib.g4go_set_parameters(self._h, theta.ctypes.data, theta.size))

    def adjoint(self, slot_weight) -> tuple[np.ndarray, dict]:
        """dLoss/dtheta of the last completed batch; slot_weight[258] = dLoss/dN + dLoss/dh_b."""
        w = np.ascontiguousarray(slot_weight, dtype=np.float64)
        if w.shape != (abi.HIST_SLOTS,):
            raise ValueError("slot_weight must hold 258 values")
        grad = np.zeros(self._n_params, np.float64)
        st = abi.AdjointStats()
        self._check(self._lib.g4go_adjoint(self._h, w.ctypes.data, grad.ctypes.data, grad.size, C.byref(st)))
        return grad, {name: getattr(st, name) for name, _ in abi.AdjointStats._fields_}

    # ---- device evaluation of the shared physics (tests) ----
    def eval_math(self, which: int, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty(2 * x.size if which == 2 else x.size, dtype=np.float32)
        self._check(self._lib.g4go_eval_math(self._h, which, x.ctypes.data, out.ctypes.data, x.size))
        return out.reshape(-1, 2) if which == 2 else out

    def eval_fresnel(self, cases: np.ndarray) -> np.ndarray:
        cases = np.ascontiguousarray(cases, dtype=np.float32).reshape(-1, 11)
        out = np.empty((cases.shape[0], 14), dtype=np.float32)
        self._check(self._lib.g4go_eval_fresnel(self._h, cases.ctypes.data, out.ctypes.data, cases.shape[0]))
        return out
