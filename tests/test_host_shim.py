"""The C ABI as a library (exports) and the C++ three-method host shim.

CPU: libg4go_hip.so loads and exports every function include/g4go_hip.h
declares; tests/cpp/shim_test checks the shim's mesh-to-box flattening, the
exception type of each failure class and an empty batch.
GPU: the shim transports scintillation, Cerenkov and direct emission records
in a crystal given as rotated/translated box meshes; its detections and
per-event counters equal the FP32 oracle's on the same records and an
equivalent box scene.  In gpu-mesh mode the crystal is a tapered mesh, which
the shim passes on as a world-space mesh solid; the oracle transports the
same mesh.
"""
import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from g4go_amd import abi
from g4go_amd.scene import (Material, Mesh, PropertyTable, Scene, Surface, SurfaceBinding, SurfaceFinish, SurfaceModel,
                            SurfaceType, Volume)

ROOT = Path(__file__).resolve().parent.parent
SHIM_TEST = ROOT / "tests" / "cpp" / "shim_test"

DETECTION_DTYPE = np.dtype([("position_mm", "<f4", 3), ("time_ns", "<f4"), ("direction", "<f4", 3),
                            ("energy_ev", "<f4"), ("event_id", "<u4"), ("photon_id", "<u4"), ("sensor_id", "<u4"),
                            ("pad", "<u4")])
STATS_DTYPE = np.dtype([("event_id", "<u4"), ("valid", "<u4"), ("captured", "<u8"), ("detected", "<u8"),
                        ("absorbed", "<u8"), ("escaped", "<u8"), ("truncated", "<u8"), ("max_bounce", "<u8"),
                        ("invalid", "<u8"), ("zero_step", "<u8")])


def declared_functions():
    text = (ROOT / "include" / "g4go_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(g4go_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_function():
    lib = abi.load()
    names = declared_functions()
    assert len(names) >= 19, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(abi.EXPORTED_SYMBOLS) <= set(names)
    lib.g4go_version.restype = ctypes.c_char_p
    assert lib.g4go_version().decode()


def shim_binary():
    if not SHIM_TEST.is_file():
        from g4go_amd import build

        build.build_shim_test(verbose=False)
    return SHIM_TEST


def test_shim_host_checks():
    proc = subprocess.run([str(shim_binary()), "cpu"], capture_output=True, text=True, timeout=60)
    assert proc.returncode == 0 and "PASS" in proc.stdout, proc.stdout + proc.stderr


REFERENCE_INCLUDE = Path(__import__("os").environ.get("G4GO_REFERENCE_ROOT", "/root/reference")) / "include"


@pytest.mark.skipif(not (REFERENCE_INCLUDE / "g4go" / "optical" / "Scene.hpp").is_file(),
                    reason="the reference tree is not present")
def test_shim_compiles_against_reference_types():
    """BasicHipTransportHost<HipTypes>, the type bundle a maintainer adds to the reference tree, type-checks
    (-fsyntax-only, every member function instantiated) against the reference's own
    PhotonTransport.hpp, PhotonTransportConfig.hpp and Scene.hpp; tests/cpp/refcheck holds a
    declaration-only G4ThreeVector, the one header the image lacks."""
    proc = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                           f"-I{ROOT / 'tests' / 'cpp' / 'refcheck'}", f"-I{REFERENCE_INCLUDE}",
                           f"-I{ROOT / 'include'}", str(ROOT / "tests" / "cpp" / "reference_types_check.cpp")],
                          capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stderr[-4000:]


def shim_crystal_scene(tapered_mesh=None):
    """The box-scene equivalent of the shim test's mesh scene (same IDs and
    tables); with tapered_mesh (world-space vertices, indices) the crystal is
    that mesh and the SiPM covers its 4x4 mm back face, as in gpu-mesh mode."""
    def t(e, v):
        return PropertyTable(np.float32(e), np.float32(v))

    sc = Scene()
    air = sc.add_material(Material(name="air", rindex=t([1.5, 4.5], [1.0, 1.0])))
    crystal = sc.add_material(Material(name="crystal", rindex=t([1.5, 4.5], [1.82, 1.82]),
                                       abs_length_mm=t([1.5, 4.5], [400.0, 400.0]),
                                       rayleigh_length_mm=t([1.5, 4.5], [500.0, 500.0]),
                                       scintillation_spectrum=[t([2.6, 2.95, 3.3], [0.0, 1.0, 0.0]), PropertyTable(),
                                                               PropertyTable()]))
    si = sc.add_material(Material(name="silicon", rindex=t([1.5, 4.5], [3.5, 3.5])))
    wrap = sc.add_surface(Surface(name="Reflector", type=SurfaceType.DielectricMetal, model=SurfaceModel.Unified,
                                  finish=SurfaceFinish.Polished, reflectivity=t([2.9], [0.95])))
    cathode = sc.add_surface(Surface(name="Cathode", type=SurfaceType.DielectricMetal, model=SurfaceModel.Unified,
                                     finish=SurfaceFinish.Polished, reflectivity=t([2.9], [0.0]),
                                     efficiency=t([2.5, 3.5], [0.3, 0.5])))
    world = sc.add_volume(Volume(name="World", material_id=air, half_mm=(100, 100, 100)))
    sc.world_volume_id = world
    mesh = None if tapered_mesh is None else Mesh(*tapered_mesh)
    xtal = sc.add_volume(Volume(name="Crystal", material_id=crystal, parent_volume_id=world, half_mm=(1.5, 1.5, 10),
                                mesh=mesh))
    sipm_half = (1.5, 1.5, 0.05) if tapered_mesh is None else (2.0, 2.0, 0.05)
    sc.add_volume(Volume(name="SiPM", material_id=si, parent_volume_id=world, skin_surface_id=cathode, sensor_id=0,
                         center_mm=(0, 0, 10.05), half_mm=sipm_half))
    sc.add_surface_binding(SurfaceBinding(xtal, world, wrap))
    return sc


def read_shim_output(path):
    raw = path.read_bytes()
    n_em, n_ev, n_det = np.frombuffer(raw, "<u8", 3)
    off = 24
    em = np.frombuffer(raw, abi.EMISSION_DTYPE, int(n_em), off)
    off += em.nbytes
    events = np.frombuffer(raw, "<u4", int(n_ev), off)
    off += events.nbytes
    idx = np.frombuffer(raw, "<u4", int(n_em), off)
    off += idx.nbytes
    det = np.frombuffer(raw, DETECTION_DTYPE, int(n_det), off)
    off += det.nbytes
    ev_stats = np.frombuffer(raw, STATS_DTYPE, int(n_ev), off)
    off += ev_stats.nbytes
    batch = np.frombuffer(raw, STATS_DTYPE, 1, off)[0]
    off += STATS_DTYPE.itemsize
    bounces = int(np.frombuffer(raw, "<u8", 1, off)[0])
    off += 8
    mesh = None
    if off < len(raw):  # gpu-mesh: the world-space mesh the shim passed on
        n_v, n_t = (int(x) for x in np.frombuffer(raw, "<u8", 2, off))
        off += 16
        verts = np.frombuffer(raw, "<f8", 3 * n_v, off).reshape(-1, 3)
        off += verts.nbytes
        mesh = (verts.copy(), np.frombuffer(raw, "<u4", 3 * n_t, off).reshape(-1, 3).copy())
    return em, events, idx, det, ev_stats, batch, bounces, mesh


def check_shim_run(tmp_path, mode):
    from oracle import oracle as O

    out = tmp_path / "shim.bin"
    proc = subprocess.run([str(shim_binary()), mode, str(out)], capture_output=True, text=True, timeout=100)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    em, events, idx, det, ev_stats, batch, bounces, mesh = read_shim_output(out)
    assert (mesh is not None) == (mode in ("gpu-mesh", "gpu-open"))
    assert set(em["type"]) == {0, 1, 2}
    ref = O.run_emissions(shim_crystal_scene(mesh), em, events, idx, 4242)
    assert len(det) == len(ref.hits) > 1000
    for name in ("position_mm", "direction", "energy_ev", "event_id", "photon_id", "sensor_id"):
        np.testing.assert_array_equal(det[name], ref.hits[name], err_msg=name)
    np.testing.assert_array_equal(det["time_ns"].view(np.uint32),
                                  ref.hits["time_ns"].astype(np.float32).view(np.uint32))
    for k, name in enumerate(("detected", "absorbed", "escaped", "truncated", "max_bounce", "invalid", "zero_step")):
        np.testing.assert_array_equal(ev_stats[name], ref.stats[:, k], err_msg=name)
        # maxBounceCount is a maximum (OptiXTransportDevice.cu:81-83), the others are sums
        expect = ref.stats[:, k].max() if name == "max_bounce" else ref.stats[:, k].sum()
        assert int(batch[name]) == int(expect), name
    np.testing.assert_array_equal(ev_stats["event_id"], events)
    np.testing.assert_array_equal(ev_stats["captured"],
                                  np.bincount(idx, weights=em["photon_count"], minlength=len(events)).astype(np.uint64))
    assert bounces == int(ref.stats[:, 7].sum())
    full = (1 << 2) | (1 << 3) | (1 << 4) | (1 << 5) | (1 << 6) | (1 << 7) | (1 << 8)
    assert int(batch["valid"]) == full | (1 << 1) | (1 << 9)
    assert all(int(v) == full | (1 << 1) for v in ev_stats["valid"])
    return mesh


@pytest.mark.gpu
def test_shim_transport_matches_oracle(tmp_path):
    print("case: C++ shim, crystal mesh scene, emission records", flush=True)
    check_shim_run(tmp_path, "gpu")


@pytest.mark.gpu
def test_shim_tapered_mesh_transport_matches_oracle(tmp_path):
    """A non-box mesh through the three-method shim: the tapered crystal, turned
    into place by its transform and passed on as a world-space mesh."""
    print("case: C++ shim, tapered crystal mesh (non-box solid), emission records", flush=True)
    verts, tris = check_shim_run(tmp_path, "gpu-mesh")
    assert len(tris) == 12 and np.isclose(verts[:, 2].max(), 10.0) and np.isclose(verts[:, 0].max(), 2.0)


@pytest.mark.gpu
def test_shim_open_mesh_transport_matches_oracle(tmp_path):
    """An open mesh through the three-method shim: the tapered crystal with
    one side face left out, an open solid the backend accepts as the
    reference does."""
    print("case: C++ shim, tapered crystal mesh with a face left out (open solid), emission records", flush=True)
    verts, tris = check_shim_run(tmp_path, "gpu-open")
    assert len(tris) == 10
