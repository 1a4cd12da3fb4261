"""MI355X (gfx950) optical-photon transport backend for G4GO.

The compute path is libg4go_hip.so (hand-written HIP kernels behind a C ABI,
include/g4go_hip.h); this package holds its Python binding, a mirror of the
reference backend interface (HipTransportHost), the workload builders and the
build recipe.
"""
from .abi import INVALID_ID, LIBRARY_PATH  # noqa: F401

__all__ = ["INVALID_ID", "LIBRARY_PATH"]
