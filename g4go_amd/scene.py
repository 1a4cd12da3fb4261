"""Scene model mirroring G4GO's backend-neutral optical scene.

Reference: include/g4go/optical/Scene.hpp:61-164 and src/optical/Scene.cpp.
A volume's solid is an axis-aligned box (centre, half lengths in mm) or a
closed convex triangle mesh in world coordinates (`Volume.mesh`, the
reference's MeshGeometry placed by its Transform), and materials may carry a
Rayleigh scattering length.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from enum import IntEnum
from typing import Optional, Sequence

import numpy as np

from . import abi

INVALID_ID = abi.INVALID_ID


class SurfaceType(IntEnum):
    DielectricDielectric = 0
    DielectricMetal = 1


class SurfaceModel(IntEnum):
    Glisur = 0
    Unified = 1


class SurfaceFinish(IntEnum):
    Polished = 0
    Ground = 1


@dataclass
class PropertyTable:
    """(energy eV, value) pairs with ascending energies (Scene.hpp:61-67)."""

    energy_ev: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    values: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))

    def __post_init__(self) -> None:
        self.energy_ev = np.ascontiguousarray(self.energy_ev, dtype=np.float32)
        self.values = np.ascontiguousarray(self.values, dtype=np.float32)
        if self.energy_ev.shape != self.values.shape or self.energy_ev.ndim != 1:
            raise ValueError("property energies and values must be 1-D arrays of equal length")

    @staticmethod
    def constant_value(value: float, energy_ev: float = 2.9) -> "PropertyTable":
        return PropertyTable(np.array([energy_ev], np.float32), np.array([value], np.float32))

    def empty(self) -> bool:
        return self.energy_ev.size == 0

    def constant(self) -> bool:
        """PropertyTable::Constant (Scene.cpp:17-21)."""
        return self.values.size == 1 or (self.values.size > 1 and bool(np.all(self.values[1:] == self.values[0])))


@dataclass
class Material:
    name: str = ""
    rindex: PropertyTable = field(default_factory=PropertyTable)
    group_velocity_mm_per_ns: PropertyTable = field(default_factory=PropertyTable)
    abs_length_mm: PropertyTable = field(default_factory=PropertyTable)
    rayleigh_length_mm: PropertyTable = field(default_factory=PropertyTable)
    scintillation_spectrum: Sequence[PropertyTable] = field(
        default_factory=lambda: [PropertyTable(), PropertyTable(), PropertyTable()])


@dataclass
class Surface:
    name: str = ""
    type: SurfaceType = SurfaceType.DielectricDielectric
    model: SurfaceModel = SurfaceModel.Unified
    finish: SurfaceFinish = SurfaceFinish.Polished
    model_value: float = 1.0
    reflectivity: PropertyTable = field(default_factory=PropertyTable)
    efficiency: PropertyTable = field(default_factory=PropertyTable)
    transmittance: PropertyTable = field(default_factory=PropertyTable)
    specular_lobe: PropertyTable = field(default_factory=PropertyTable)
    specular_spike: PropertyTable = field(default_factory=PropertyTable)
    backscatter: PropertyTable = field(default_factory=PropertyTable)
    surface_roughness: PropertyTable = field(default_factory=PropertyTable)


@dataclass
class Mesh:
    """Closed convex triangle mesh in world coordinates (MeshGeometry, Scene.hpp:94-100).

    vertices_mm: (n, 3) float64; indices: (m, 3) uint32; triangle_normals:
    optional (m, 3) float64 boundary normals (else each triangle's plane
    normal).  The winding does not matter.
    """
    vertices_mm: np.ndarray
    indices: np.ndarray
    triangle_normals: Optional[np.ndarray] = None

    def __post_init__(self) -> None:
        self.vertices_mm = np.ascontiguousarray(self.vertices_mm, dtype=np.float64).reshape(-1, 3)
        self.indices = np.ascontiguousarray(self.indices, dtype=np.uint32).reshape(-1, 3)
        if self.triangle_normals is not None:
            self.triangle_normals = np.ascontiguousarray(self.triangle_normals, dtype=np.float64).reshape(-1, 3)

    def transformed(self, rotation=None, translation_mm=(0.0, 0.0, 0.0)) -> "Mesh":
        """The mesh placed by a Transform (Scene.hpp:42-59): x -> R x + t."""
        r = np.eye(3) if rotation is None else np.asarray(rotation, dtype=np.float64).reshape(3, 3)
        t = np.asarray(translation_mm, dtype=np.float64)
        normals = None if self.triangle_normals is None else self.triangle_normals @ r.T
        return Mesh(self.vertices_mm @ r.T + t, self.indices.copy(), normals)


def frustum_mesh(front_half_xy: Sequence[float], back_half_xy: Sequence[float], half_z: float) -> Mesh:
    """Frustum of a rectangular pyramid, 12 triangles (two per face): the
    rectangle front_half_xy at z = -half_z, back_half_xy at z = +half_z."""
    (fx, fy), (bx, by), hz = (float(x) for x in front_half_xy), (float(x) for x in back_half_xy), float(half_z)
    v = np.array([[sx * (fx if z < 0 else bx), sy * (fy if z < 0 else by), z]
                  for z in (-hz, hz) for sy in (-1.0, 1.0) for sx in (-1.0, 1.0)])
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    tri = [(a, b, c) for a, b, c, d in quads] + [(a, c, d) for a, b, c, d in quads]
    return Mesh(v, np.array(tri))


def box_mesh(half_mm: Sequence[float]) -> Mesh:
    """The 12-triangle mesh of the box [-h, h] (two triangles per face)."""
    return frustum_mesh(half_mm[:2], half_mm[:2], half_mm[2])


def _ear_clip(poly: np.ndarray) -> list:
    """Triangles (counter-clockwise index triples) of a simple counter-clockwise polygon."""
    def cross(o, a, b):
        return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])

    idx = list(range(len(poly)))
    tris = []
    while len(idx) > 3:
        for k in range(len(idx)):
            i0, i1, i2 = idx[k - 1], idx[k], idx[(k + 1) % len(idx)]
            p0, p1, p2 = poly[i0], poly[i1], poly[i2]
            if cross(p0, p1, p2) <= 0:
                continue  # reflex corner
            if any(cross(p0, p1, poly[j]) >= 0 and cross(p1, p2, poly[j]) >= 0 and cross(p2, p0, poly[j]) >= 0
                   for j in idx if j not in (i0, i1, i2)):
                continue  # another vertex inside the ear
            tris.append((i0, i1, i2))
            idx.pop(k)
            break
        else:
            raise ValueError("polygon is not simple and counter-clockwise")
    tris.append(tuple(idx))
    return tris


def prism_mesh(polygon_xy: Sequence[Sequence[float]], z_lo: float, z_hi: float) -> Mesh:
    """Right prism over a simple polygon (counter-clockwise xy vertices, convex
    or not) between z_lo and z_hi, wound outwards; caps by ear clipping."""
    poly = np.asarray(polygon_xy, dtype=np.float64)
    n = len(poly)
    v = np.concatenate([np.c_[poly, np.full(n, z_lo)], np.c_[poly, np.full(n, z_hi)]])
    tri = []
    for a, b, c in _ear_clip(poly):
        tri.append((a, c, b))              # bottom, facing -z
        tri.append((n + a, n + b, n + c))  # top, facing +z
    for i in range(n):
        j = (i + 1) % n
        tri.append((i, j, n + j))
        tri.append((i, n + j, n + i))
    return Mesh(v, np.array(tri))


def tube_mesh(r_in: float, r_out: float, half_z: float, segments: int = 24) -> Mesh:
    """Hollow cylinder (a non-convex solid) about the z axis: regular polygons
    of `segments` sides inscribed in r_in and r_out, |z| <= half_z, wound outwards."""
    ang = 2.0 * np.pi * np.arange(segments) / segments
    ring = np.c_[np.cos(ang), np.sin(ang)]
    v = np.concatenate([np.c_[r_out * ring, np.full(segments, -half_z)], np.c_[r_out * ring, np.full(segments, half_z)],
                        np.c_[r_in * ring, np.full(segments, -half_z)], np.c_[r_in * ring, np.full(segments, half_z)]])
    ob, ot, ib, it = 0, segments, 2 * segments, 3 * segments
    tri = []
    for k in range(segments):
        m = (k + 1) % segments
        tri += [(ob + k, ob + m, ot + m), (ob + k, ot + m, ot + k)]   # outer wall, facing out
        tri += [(ib + m, ib + k, it + k), (ib + m, it + k, it + m)]   # inner wall, facing the axis
        tri += [(ob + k, ib + k, ib + m), (ob + k, ib + m, ob + m)]   # bottom annulus, facing -z
        tri += [(ot + k, ot + m, it + m), (ot + k, it + m, it + k)]   # top annulus, facing +z
    return Mesh(v, np.array(tri))


@dataclass
class Volume:
    name: str = ""
    volume_id: int = INVALID_ID
    material_id: int = INVALID_ID
    parent_volume_id: int = INVALID_ID
    skin_surface_id: int = INVALID_ID
    sensor_id: int = INVALID_ID
    center_mm: Sequence[float] = (0.0, 0.0, 0.0)
    half_mm: Sequence[float] = (1.0, 1.0, 1.0)
    mesh: Optional[Mesh] = None  # the solid when set (centre and half lengths are then ignored)


@dataclass
class SurfaceBinding:
    from_volume_id: int = INVALID_ID
    to_volume_id: int = INVALID_ID
    surface_id: int = INVALID_ID


class Scene:
    """Scene (Scene.hpp:128-164) with the same add/find API."""

    def __init__(self) -> None:
        self.materials: list[Material] = []
        self.surfaces: list[Surface] = []
        self.volumes: list[Volume] = []
        self.surface_bindings: list[SurfaceBinding] = []
        self.world_volume_id: int = INVALID_ID

    def add_material(self, material: Material) -> int:
        self.materials.append(material)
        return len(self.materials) - 1

    def add_surface(self, surface: Surface) -> int:
        self.surfaces.append(surface)
        return len(self.surfaces) - 1

    def add_volume(self, volume: Volume) -> int:
        """Scene::AddVolume: an unset ID becomes the volume's index (Scene.cpp:39-45)."""
        if volume.volume_id == INVALID_ID:
            volume.volume_id = len(self.volumes)
        self.volumes.append(volume)
        return len(self.volumes) - 1

    def add_surface_binding(self, binding: SurfaceBinding) -> None:
        self.surface_bindings.append(binding)

    def find_surface(self, surface_id: int) -> Optional[Surface]:
        return self.surfaces[surface_id] if 0 <= surface_id < len(self.surfaces) else None

    def find_volume(self, volume_id: int) -> Optional[Volume]:
        if 0 <= volume_id < len(self.volumes) and self.volumes[volume_id].volume_id == volume_id:
            return self.volumes[volume_id]
        return next((v for v in self.volumes if v.volume_id == volume_id), None)

    def find_boundary_surface(self, from_volume_id: int, to_volume_id: int) -> Optional[Surface]:
        """Scene::FindBoundarySurface (Scene.cpp:120-150): border binding, else skin precedence."""
        for b in self.surface_bindings:
            if b.from_volume_id == from_volume_id and b.to_volume_id == to_volume_id:
                return self.find_surface(b.surface_id)
        src = self.find_volume(from_volume_id)
        dst = self.find_volume(to_volume_id)
        if src is None or dst is None:
            return None

        def skin(v: Volume) -> Optional[Surface]:
            return None if v.skin_surface_id == INVALID_ID else self.find_surface(v.skin_surface_id)

        if dst.parent_volume_id == from_volume_id:
            return skin(dst) or skin(src)
        return skin(src) or skin(dst)

    # ---- C ABI view ----
    def to_abi(self) -> tuple[abi.Scene, list]:
        """Returns the g4go_scene description and the objects that keep its arrays alive."""
        keep: list = []

        def prop(p: PropertyTable) -> abi.Property:
            if p.empty():
                return abi.Property(0, None, None)
            keep.extend([p.energy_ev, p.values])
            return abi.Property(p.energy_ev.size, p.energy_ev.ctypes.data_as(C.POINTER(C.c_float)),
                                p.values.ctypes.data_as(C.POINTER(C.c_float)))

        mats = (abi.Material * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            mats[i].rindex = prop(m.rindex)
            mats[i].group_velocity_mm_per_ns = prop(m.group_velocity_mm_per_ns)
            mats[i].abs_length_mm = prop(m.abs_length_mm)
            mats[i].rayleigh_length_mm = prop(m.rayleigh_length_mm)
            for k in range(3):
                mats[i].scintillation_spectrum[k] = prop(m.scintillation_spectrum[k])
        surfs = (abi.Surface * max(1, len(self.surfaces)))()
        for i, s in enumerate(self.surfaces):
            surfs[i].type = int(s.type)
            surfs[i].model = int(s.model)
            surfs[i].finish = int(s.finish)
            surfs[i].model_value = float(s.model_value)
            for name in ("reflectivity", "efficiency", "transmittance", "specular_lobe", "specular_spike",
                         "backscatter", "surface_roughness"):
                setattr(surfs[i], name, prop(getattr(s, name)))
        vols = (abi.Volume * max(1, len(self.volumes)))()
        for i, v in enumerate(self.volumes):
            vols[i].volume_id = v.volume_id
            vols[i].material_id = v.material_id
            vols[i].parent_volume_id = v.parent_volume_id
            vols[i].skin_surface_id = v.skin_surface_id
            vols[i].sensor_id = v.sensor_id
            vols[i].center_mm[:] = [float(x) for x in v.center_mm]
            vols[i].half_mm[:] = [float(x) for x in v.half_mm]
        binds = (abi.SurfaceBinding * max(1, len(self.surface_bindings)))()
        for i, b in enumerate(self.surface_bindings):
            binds[i].from_volume_id = b.from_volume_id
            binds[i].to_volume_id = b.to_volume_id
            binds[i].surface_id = b.surface_id
        meshed = [v for v in self.volumes if v.mesh is not None]
        meshes = (abi.VolumeMesh * max(1, len(meshed)))()
        for i, v in enumerate(meshed):
            m = v.mesh
            keep.extend([m.vertices_mm, m.indices, m.triangle_normals])
            meshes[i].volume_id = v.volume_id
            meshes[i].n_vertices = len(m.vertices_mm)
            meshes[i].vertices_mm = m.vertices_mm.ctypes.data_as(C.POINTER(C.c_double))
            meshes[i].indices = m.indices.ctypes.data_as(C.POINTER(C.c_uint32))
            meshes[i].n_triangles = len(m.indices)
            if m.triangle_normals is not None:
                meshes[i].triangle_normals = m.triangle_normals.ctypes.data_as(C.POINTER(C.c_double))
        keep.extend([mats, surfs, vols, binds, meshes])
        desc = abi.Scene(mats, len(self.materials), surfs, len(self.surfaces), vols, len(self.volumes), binds,
                         len(self.surface_bindings), self.world_volume_id, meshes, len(meshed))
        return desc, keep
