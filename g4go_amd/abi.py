"""ctypes mirror of include/g4go_hip.h and the loader of libg4go_hip.so.

The library is loaded from the package directory only (built in-tree by
g4go_amd.build); a missing library raises instead of falling back to any
other implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIBRARY_PATH = Path(__file__).resolve().parent / "libg4go_hip.so"
INVALID_ID = 0xFFFFFFFF
HIST_BINS = 256
HIST_SLOTS = 258

G4GO_OK = 0
G4GO_ERROR_RUNTIME = 1
G4GO_ERROR_INVALID_ARGUMENT = 2
G4GO_ERROR_OVERFLOW = 3
G4GO_ERROR_LOGIC = 4


class Config(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("max_photons_per_event", C.c_uint32),
        ("max_bounces_per_photon", C.c_uint32),
        ("boundary_tolerance_mm", C.c_float),
        ("target_photons_per_batch", C.c_uint32),
        ("max_events_per_batch", C.c_uint32),
        ("enable_performance_diagnostics", C.c_uint32),
        ("device", C.c_int32),
        ("sort_hits", C.c_uint32),
        ("max_in_flight", C.c_uint32),
    ]


class Launch(C.Structure):
    _fields_ = [
        ("persistent", C.c_uint32),
        ("block_size", C.c_uint32),
        ("chunk", C.c_uint32),
        ("refill_threshold", C.c_uint32),
        ("blocks_per_cu", C.c_uint32),
        ("per_event", C.c_uint32),
        ("graph", C.c_uint32),
        ("pregenerate", C.c_uint32),
    ]


class Property(C.Structure):
    _fields_ = [("count", C.c_uint32), ("energy_ev", C.POINTER(C.c_float)), ("values", C.POINTER(C.c_float))]


class Material(C.Structure):
    _fields_ = [
        ("rindex", Property),
        ("group_velocity_mm_per_ns", Property),
        ("abs_length_mm", Property),
        ("rayleigh_length_mm", Property),
        ("scintillation_spectrum", Property * 3),
    ]


class Surface(C.Structure):
    _fields_ = [
        ("type", C.c_uint32),
        ("model", C.c_uint32),
        ("finish", C.c_uint32),
        ("model_value", C.c_float),
        ("reflectivity", Property),
        ("efficiency", Property),
        ("transmittance", Property),
        ("specular_lobe", Property),
        ("specular_spike", Property),
        ("backscatter", Property),
        ("surface_roughness", Property),
    ]


class Volume(C.Structure):
    _fields_ = [
        ("volume_id", C.c_uint32),
        ("material_id", C.c_uint32),
        ("parent_volume_id", C.c_uint32),
        ("skin_surface_id", C.c_uint32),
        ("sensor_id", C.c_uint32),
        ("center_mm", C.c_float * 3),
        ("half_mm", C.c_float * 3),
    ]


class SurfaceBinding(C.Structure):
    _fields_ = [("from_volume_id", C.c_uint32), ("to_volume_id", C.c_uint32), ("surface_id", C.c_uint32)]


class VolumeMesh(C.Structure):
    _fields_ = [
        ("volume_id", C.c_uint32),
        ("n_vertices", C.c_uint32),
        ("vertices_mm", C.POINTER(C.c_double)),
        ("indices", C.POINTER(C.c_uint32)),
        ("n_triangles", C.c_uint32),
        ("reserved", C.c_uint32),
        ("triangle_normals", C.POINTER(C.c_double)),
    ]


class Scene(C.Structure):
    _fields_ = [
        ("materials", C.POINTER(Material)),
        ("n_materials", C.c_uint32),
        ("surfaces", C.POINTER(Surface)),
        ("n_surfaces", C.c_uint32),
        ("volumes", C.POINTER(Volume)),
        ("n_volumes", C.c_uint32),
        ("bindings", C.POINTER(SurfaceBinding)),
        ("n_bindings", C.c_uint32),
        ("world_volume_id", C.c_uint32),
        ("meshes", C.POINTER(VolumeMesh)),
        ("n_meshes", C.c_uint32),
    ]


class Emission(C.Structure):
    _fields_ = [
        ("position_mm", C.c_float * 3),
        ("time_ns", C.c_float),
        ("direction", C.c_float * 3),
        ("step_length_mm", C.c_float),
        ("step_delta_mm", C.c_float * 3),
        ("pre_velocity_mm_per_ns", C.c_float),
        ("polarization", C.c_float * 3),
        ("delta_velocity_mm_per_ns", C.c_float),
        ("energy_ev", C.c_float),
        ("pre_mean_photon_count", C.c_float),
        ("post_mean_photon_count", C.c_float),
        ("decay_time_ns", C.c_float),
        ("rise_time_ns", C.c_float),
        ("charge", C.c_float),
        ("event_id", C.c_uint32),
        ("first_photon_id", C.c_uint32),
        ("photon_count", C.c_uint32),
        ("volume_id", C.c_uint32),
        ("material_id", C.c_uint32),
        ("spectrum_id", C.c_uint32),
        ("type", C.c_uint8),
        ("reserved", C.c_uint8 * 15),
    ]


class SyntheticSource(C.Structure):
    _fields_ = [
        ("lo_mm", C.c_float * 3),
        ("hi_mm", C.c_float * 3),
        ("wavelength_mean_nm", C.c_float),
        ("wavelength_sigma_nm", C.c_float),
        ("decay_time_ns", C.c_float),
        ("volume_id", C.c_uint32),
    ]


class ParamTable(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("index", C.c_uint32)]


class AdjointStats(C.Structure):
    _fields_ = [("replayed", C.c_uint64), ("mismatched", C.c_uint64), ("rescored", C.c_uint64),
                ("steps", C.c_uint64), ("kernel_ms", C.c_double)]


PARAM_ABS_LENGTH, PARAM_RAYLEIGH_LENGTH, PARAM_REFLECTIVITY, PARAM_EFFICIENCY = 0, 1, 2, 3


class Hit(C.Structure):
    _fields_ = [
        ("position_mm", C.c_float * 3),
        ("direction", C.c_float * 3),
        ("energy_ev", C.c_float),
        ("sensor_id", C.c_uint32),
        ("event_id", C.c_uint32),
        ("photon_id", C.c_uint32),
        ("time_ns", C.c_double),
    ]


class EventStats(C.Structure):
    _fields_ = [
        ("event_id", C.c_uint32),
        ("reserved", C.c_uint32),
        ("captured", C.c_uint64),
        ("detected", C.c_uint64),
        ("absorbed", C.c_uint64),
        ("escaped", C.c_uint64),
        ("truncated", C.c_uint64),
        ("max_bounce", C.c_uint64),
        ("invalid", C.c_uint64),
        ("zero_step", C.c_uint64),
    ]


class BatchStats(C.Structure):
    _fields_ = [
        ("captured", C.c_uint64),
        ("generated", C.c_uint64),
        ("detected", C.c_uint64),
        ("absorbed", C.c_uint64),
        ("escaped", C.c_uint64),
        ("truncated", C.c_uint64),
        ("max_bounce", C.c_uint64),
        ("invalid", C.c_uint64),
        ("zero_step", C.c_uint64),
        ("total_steps", C.c_uint64),
        ("dropped_hits", C.c_uint64),
        ("transport_time_ms", C.c_double),
        ("kernel_ms", C.c_double),
        ("host_to_device_ms", C.c_double),
        ("device_to_host_ms", C.c_double),
        ("host_to_device_bytes", C.c_uint64),
        ("device_to_host_bytes", C.c_uint64),
        ("device_memset_ms", C.c_double),
        ("device_order_ms", C.c_double),
        ("metadata_to_host_ms", C.c_double),
        ("metadata_to_host_bytes", C.c_uint64),
        ("hits_to_host_ms", C.c_double),
        ("hits_to_host_bytes", C.c_uint64),
    ]


class BatchResult(C.Structure):
    _fields_ = [
        ("hits", C.POINTER(Hit)),
        ("n_hits", C.c_uint64),
        ("events", C.POINTER(EventStats)),
        ("n_events", C.c_uint64),
        ("histogram", C.POINTER(C.c_uint64)),
        ("stats", BatchStats),
    ]


NOALLOC_MAX_EVENTS = 1024
NOALLOC_SCRATCH_BYTES = 1844224


class DeviceOutputs(C.Structure):
    """g4go_device_outputs: device pointers of an allocation-free batch's outputs."""
    _fields_ = [
        ("hits", C.c_void_p),
        ("hit_capacity", C.c_uint64),
        ("hit_count", C.c_void_p),
        ("histogram", C.c_void_p),
        ("event_counters", C.c_void_p),
        ("batch_counters", C.c_void_p),
    ]


HIT_DTYPE = np.dtype([
    ("position_mm", "<f4", (3,)),
    ("direction", "<f4", (3,)),
    ("energy_ev", "<f4"),
    ("sensor_id", "<u4"),
    ("event_id", "<u4"),
    ("photon_id", "<u4"),
    ("time_ns", "<f8"),
])
assert HIT_DTYPE.itemsize == C.sizeof(Hit) == 48

EVENT_DTYPE = np.dtype([
    ("event_id", "<u4"), ("reserved", "<u4"), ("captured", "<u8"), ("detected", "<u8"), ("absorbed", "<u8"),
    ("escaped", "<u8"), ("truncated", "<u8"), ("max_bounce", "<u8"), ("invalid", "<u8"), ("zero_step", "<u8"),
])
assert EVENT_DTYPE.itemsize == C.sizeof(EventStats)

EMISSION_DTYPE = np.dtype([
    ("position_mm", "<f4", (3,)), ("time_ns", "<f4"), ("direction", "<f4", (3,)), ("step_length_mm", "<f4"),
    ("step_delta_mm", "<f4", (3,)), ("pre_velocity_mm_per_ns", "<f4"), ("polarization", "<f4", (3,)),
    ("delta_velocity_mm_per_ns", "<f4"), ("energy_ev", "<f4"), ("pre_mean_photon_count", "<f4"),
    ("post_mean_photon_count", "<f4"), ("decay_time_ns", "<f4"), ("rise_time_ns", "<f4"), ("charge", "<f4"),
    ("event_id", "<u4"), ("first_photon_id", "<u4"), ("photon_count", "<u4"), ("volume_id", "<u4"),
    ("material_id", "<u4"), ("spectrum_id", "<u4"), ("type", "u1"), ("reserved", "u1", (15,)),
])
assert EMISSION_DTYPE.itemsize == C.sizeof(Emission) == 128

assert C.sizeof(Config) == 48 and C.sizeof(Launch) == 32

EXPORTED_SYMBOLS = (
    "g4go_create", "g4go_destroy", "g4go_last_error", "g4go_set_launch", "g4go_get_launch",
    "g4go_prepare_scene", "g4go_enqueue_emissions", "g4go_enqueue_synthetic", "g4go_complete",
    "g4go_complete_device", "g4go_eval_math", "g4go_eval_fresnel", "g4go_device_info", "g4go_trace_synthetic", "g4go_set_parameter_layout", "g4go_get_parameters",
    "g4go_set_parameters", "g4go_adjoint", "g4go_version", "g4go_set_stream", "g4go_enqueue_emissions_device",
    "g4go_transport_synthetic_noalloc",
)

_lib = None


def load() -> C.CDLL:
    """Loads the in-tree HIP library, failing loudly when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIBRARY_PATH.is_file():
        raise RuntimeError(f"{LIBRARY_PATH} is missing: build it with `python -m g4go_amd.build`")
    # torch ships its own HIP runtime under the same soname (libamdhip64.so.7)
    # as the system one this library links.  Loaded first, torch's copy serves
    # both; loaded after a process has initialised the system copy, torch
    # finds no device.  So torch is imported before the library.
    import torch  # noqa: F401
    lib = C.CDLL(str(LIBRARY_PATH))
    H = C.c_void_p
    lib.g4go_create.argtypes = [C.POINTER(Config), C.POINTER(H)]
    lib.g4go_destroy.argtypes = [H]
    lib.g4go_destroy.restype = None
    lib.g4go_last_error.argtypes = [H]
    lib.g4go_last_error.restype = C.c_char_p
    lib.g4go_set_launch.argtypes = [H, C.POINTER(Launch)]
    lib.g4go_get_launch.argtypes = [H, C.POINTER(Launch)]
    lib.g4go_prepare_scene.argtypes = [H, C.POINTER(Scene)]
    lib.g4go_enqueue_emissions.argtypes = [H, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p]
    lib.g4go_enqueue_emissions_device.argtypes = [H, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p]
    lib.g4go_set_stream.argtypes = [H, C.c_void_p]
    lib.g4go_enqueue_synthetic.argtypes = [H, C.POINTER(SyntheticSource), C.c_void_p, C.c_void_p, C.c_uint64]
    lib.g4go_complete.argtypes = [H, C.POINTER(BatchResult)]
    lib.g4go_complete_device.argtypes = [H, C.POINTER(BatchResult), C.POINTER(C.c_void_p)]
    lib.g4go_eval_math.argtypes = [H, C.c_int, C.c_void_p, C.c_void_p, C.c_uint64]
    lib.g4go_eval_fresnel.argtypes = [H, C.c_void_p, C.c_void_p, C.c_uint64]
    lib.g4go_device_info.argtypes = [H, C.c_char_p, C.c_uint64, C.POINTER(C.c_int32)]
    lib.g4go_trace_synthetic.argtypes = [H, C.POINTER(SyntheticSource), C.c_uint32, C.c_uint32, C.c_uint32,
                                         C.c_void_p, C.POINTER(C.c_uint32)]
    lib.g4go_set_parameter_layout.argtypes = [H, C.POINTER(ParamTable), C.c_uint32, C.POINTER(C.c_uint32)]
    lib.g4go_get_parameters.argtypes = [H, C.c_void_p, C.c_uint64]
    lib.g4go_set_parameters.argtypes = [H, C.c_void_p, C.c_uint64]
    lib.g4go_adjoint.argtypes = [H, C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(AdjointStats)]
    lib.g4go_transport_synthetic_noalloc.argtypes = [H, C.POINTER(SyntheticSource), C.c_void_p, C.c_void_p,
                                                     C.c_uint32, C.POINTER(DeviceOutputs), C.c_void_p]
    lib.g4go_version.restype = C.c_char_p
    for name in ("g4go_create", "g4go_set_launch", "g4go_get_launch", "g4go_prepare_scene", "g4go_enqueue_emissions",
                 "g4go_enqueue_synthetic", "g4go_complete", "g4go_complete_device", "g4go_eval_math", "g4go_eval_fresnel", "g4go_device_info", "g4go_trace_synthetic",
                 "g4go_set_parameter_layout", "g4go_get_parameters", "g4go_set_parameters", "g4go_adjoint",
                 "g4go_set_stream", "g4go_enqueue_emissions_device", "g4go_transport_synthetic_noalloc"):
        getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib
