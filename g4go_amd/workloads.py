"""Scenes and seeded inputs of the five benchmark workloads.

* single crystal 3x3x20 mm (n = 1.82, absorption 400 mm in 32 wavelength
  nodes, Rayleigh 500 mm) in an air world, wrapped on five faces by a
  polished dielectric_metal reflector (R = 0.95, the reference's
  "Reflector" border surface, DetectorConstruction.cpp:149-154) and read out
  by a 3x3 mm SiPM on the +z face (polished dielectric_metal "Cathode" skin
  with R = 0 and a 32-node PDE curve resampled from the reference's sensor
  efficiency table, DetectorConstruction.cpp:62-68, 163-169);
* an 8x8 array of the same crystals at 3.2 mm pitch with reflective septa
  (one wrap border surface per crystal) and one SiPM per crystal;
* synthetic photons uniform in the crystal (isotropic, random linear
  polarisation, wavelength N(420, 20) nm, emission time Exp(40 ns));
* emission records of 1 MeV electrons on the crystal axis, synthesised by a
  seeded generator (Geant4 is not available to produce them);
* log-normal event sizes for the batched workloads;
* mesh solids: the crystal tapered from a 3x3 mm front face to a 4x4 mm back
  face (a frustum given as a 12-triangle convex mesh, SiPM box on the back
  face), the single-crystal scene turned by a rotation (crystal and SiPM as
  box meshes), an L-shaped light guide (a non-convex prism, 20 triangles)
  read out at the end of its long arm, and a ring of 48 crystals turned to
  face its centre (box meshes), read out on their outer faces.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import abi
from .scene import (INVALID_ID, Material, PropertyTable, Scene, Surface, SurfaceBinding, SurfaceFinish,
                    SurfaceModel, SurfaceType, Volume, box_mesh, frustum_mesh, prism_mesh)

HC_EV_NM = 1239.84198
N_NODES = 32
WAVELENGTH_LO_NM = 340.0
WAVELENGTH_HI_NM = 500.0

# Reference SiPM efficiency curve (DetectorConstruction.cpp:62-68), energy in eV.
REFERENCE_SENSOR_ENERGY_EV = np.array([1.771068, 2.101763, 2.478823, 2.66882, 2.756137, 2.849362, 2.949113, 3.02425,
                                       3.099632, 3.178868, 3.3981, 3.649811, 3.877416, 4.128759, 4.348776])
REFERENCE_SENSOR_EFFICIENCY = np.array([0.26319, 0.395706, 0.503681, 0.525767, 0.555215, 0.570552, 0.597546,
                                        0.59816, 0.599387, 0.593252, 0.521472, 0.466258, 0.458896, 0.408589,
                                        0.196933])


def node_energies() -> np.ndarray:
    """Energies (ascending) of the 32 wavelength-bin centres over 340-500 nm (5 nm bins)."""
    width = (WAVELENGTH_HI_NM - WAVELENGTH_LO_NM) / N_NODES
    wl = WAVELENGTH_LO_NM + width * (np.arange(N_NODES) + 0.5)
    return np.sort((HC_EV_NM / wl).astype(np.float32))


def reference_pde_nodes() -> np.ndarray:
    """Reference SiPM efficiency, linearly interpolated at the node energies."""
    return np.interp(node_energies().astype(np.float64), REFERENCE_SENSOR_ENERGY_EV,
                     REFERENCE_SENSOR_EFFICIENCY).astype(np.float32)


def gaussian_wavelength_spectrum(mean_nm: float = 420.0, sigma_nm: float = 20.0, points: int = 64) -> PropertyTable:
    """Emission spectrum N(mean, sigma) in wavelength, as a density over energy."""
    wl = np.linspace(mean_nm - 4 * sigma_nm, mean_nm + 4 * sigma_nm, points)
    e = HC_EV_NM / wl
    density = np.exp(-0.5 * ((wl - mean_nm) / sigma_nm) ** 2) * wl * wl / HC_EV_NM
    order = np.argsort(e)
    return PropertyTable(e[order].astype(np.float32), density[order].astype(np.float32))


@dataclass
class CrystalParams:
    rindex: float = 1.82
    abs_length_mm: float = 400.0
    rayleigh_length_mm: float = 500.0
    wrap_reflectivity: float = 0.95
    half_mm: tuple = (1.5, 1.5, 10.0)
    sipm_thickness_mm: float = 0.1


def _crystal_material(p: CrystalParams, abs_nodes=None, name="crystal") -> Material:
    nodes = node_energies()
    absl = np.full(N_NODES, p.abs_length_mm, np.float32) if abs_nodes is None else np.asarray(abs_nodes, np.float32)
    return Material(
        name=name,
        rindex=PropertyTable.constant_value(p.rindex),
        abs_length_mm=PropertyTable(nodes, absl),
        rayleigh_length_mm=PropertyTable.constant_value(p.rayleigh_length_mm),
        scintillation_spectrum=[gaussian_wavelength_spectrum(), PropertyTable(), PropertyTable()],
    )


def _common_materials_and_surfaces(scene: Scene, p: CrystalParams, pde_nodes=None):
    air = scene.add_material(Material(name="air", rindex=PropertyTable.constant_value(1.0)))
    silicon = scene.add_material(Material(name="silicon", rindex=PropertyTable.constant_value(3.5)))
    wrap = scene.add_surface(Surface(name="Reflector", type=SurfaceType.DielectricMetal, model=SurfaceModel.Unified,
                                     finish=SurfaceFinish.Polished,
                                     reflectivity=PropertyTable.constant_value(p.wrap_reflectivity)))
    pde = reference_pde_nodes() if pde_nodes is None else np.asarray(pde_nodes, np.float32)
    cathode = scene.add_surface(Surface(name="Cathode", type=SurfaceType.DielectricMetal, model=SurfaceModel.Unified,
                                        finish=SurfaceFinish.Polished,
                                        reflectivity=PropertyTable.constant_value(0.0),
                                        efficiency=PropertyTable(node_energies(), pde)))
    return air, silicon, wrap, cathode


@dataclass
class CrystalScene:
    scene: Scene
    world_id: int
    crystal_ids: list
    sipm_ids: list
    crystal_materials: list
    wrap_surface: int
    cathode_surface: int
    params: CrystalParams


def crystal_scene(p: CrystalParams | None = None, abs_nodes=None, pde_nodes=None) -> CrystalScene:
    """Single wrapped crystal with a SiPM on its +z face."""
    p = p or CrystalParams()
    scene = Scene()
    air, silicon, wrap, cathode = _common_materials_and_surfaces(scene, p, pde_nodes)
    crystal_mat = scene.add_material(_crystal_material(p, abs_nodes))
    world = scene.add_volume(Volume(name="World", material_id=air, center_mm=(0, 0, 0), half_mm=(100, 100, 100)))
    scene.world_volume_id = world
    hz = p.half_mm[2]
    crystal = scene.add_volume(Volume(name="Crystal", material_id=crystal_mat, parent_volume_id=world,
                                      center_mm=(0, 0, 0), half_mm=p.half_mm))
    sipm = scene.add_volume(Volume(name="SiPM", material_id=silicon, parent_volume_id=world,
                                   skin_surface_id=cathode, sensor_id=0,
                                   center_mm=(0, 0, hz + 0.5 * p.sipm_thickness_mm),
                                   half_mm=(p.half_mm[0], p.half_mm[1], 0.5 * p.sipm_thickness_mm)))
    scene.add_surface_binding(SurfaceBinding(crystal, world, wrap))
    return CrystalScene(scene, world, [crystal], [sipm], [crystal_mat], wrap, cathode, p)


GROUND_FINISHES = ("glisur_metal", "unified_metal", "glisur_dielectric", "unified_dielectric")


def ground_wrap(kind: str) -> Surface:
    """The crystal's border surface with a ground finish (OptiXTransportDevice.cu:492-595, 843-876):
    * glisur_metal: dielectric_metal reflector, R = 0.95, glisur polish 0.7 (smeared facets);
    * unified_metal: dielectric_metal reflector, R = 0.95, unified sigma_alpha 0.2 rad with
      spike 0.2 / lobe 0.5 / backscatter 0.1 (Lambertian 0.2);
    * glisur_dielectric: bare ground crystal/air interface, glisur polish 0.8, surface roughness
      20 nm (Lambertian reflections from the dense side);
    * unified_dielectric: bare ground interface, unified sigma_alpha 0.15 rad, spike 0.3 / lobe 0.5 /
      backscatter 0.05, roughness 20 nm.
    """
    c = PropertyTable.constant_value
    if kind == "glisur_metal":
        return Surface(name="GroundGlisur", type=SurfaceType.DielectricMetal, model=SurfaceModel.Glisur,
                       finish=SurfaceFinish.Ground, model_value=0.7, reflectivity=c(0.95))
    if kind == "unified_metal":
        return Surface(name="GroundUnified", type=SurfaceType.DielectricMetal, model=SurfaceModel.Unified,
                       finish=SurfaceFinish.Ground, model_value=0.2, reflectivity=c(0.95),
                       specular_spike=c(0.2), specular_lobe=c(0.5), backscatter=c(0.1))
    if kind == "glisur_dielectric":
        return Surface(name="GroundGlisurBare", type=SurfaceType.DielectricDielectric, model=SurfaceModel.Glisur,
                       finish=SurfaceFinish.Ground, model_value=0.8, surface_roughness=c(2.0e-5))
    if kind == "unified_dielectric":
        return Surface(name="GroundUnifiedBare", type=SurfaceType.DielectricDielectric, model=SurfaceModel.Unified,
                       finish=SurfaceFinish.Ground, model_value=0.15, specular_spike=c(0.3),
                       specular_lobe=c(0.5), backscatter=c(0.05), surface_roughness=c(2.0e-5))
    raise ValueError(f"unknown ground finish {kind!r}")


def ground_crystal_scene(kind: str, p: CrystalParams | None = None) -> CrystalScene:
    """crystal_scene with its polished wrap replaced by the ground border surface ground_wrap(kind)."""
    cs = crystal_scene(p)
    cs.scene.surfaces[cs.wrap_surface] = ground_wrap(kind)
    return cs


def array_scene(p: CrystalParams | None = None, n: int = 8, pitch_mm: float = 3.2, abs_nodes=None,
                pde_nodes=None, septa: bool = True) -> CrystalScene:
    """n x n crystals with reflective septa, one SiPM per crystal (2 n^2 boxes).

    septa=False leaves the crystal faces bare (Fresnel crystal/air), so photons
    cross the world between crystals.
    """
    p = p or CrystalParams()
    scene = Scene()
    air, silicon, wrap, cathode = _common_materials_and_surfaces(scene, p, pde_nodes)
    world = scene.add_volume(Volume(name="World", material_id=air, center_mm=(0, 0, 0), half_mm=(100, 100, 100)))
    scene.world_volume_id = world
    hz = p.half_mm[2]
    crystals, sipms, mats = [], [], []
    for i in range(n):
        for j in range(n):
            k = i * n + j
            x = (i - 0.5 * (n - 1)) * pitch_mm
            y = (j - 0.5 * (n - 1)) * pitch_mm
            nodes = None if abs_nodes is None else np.asarray(abs_nodes)[k]
            m = scene.add_material(_crystal_material(p, nodes, name=f"crystal{k}"))
            c = scene.add_volume(Volume(name=f"Crystal{k}", material_id=m, parent_volume_id=world,
                                        center_mm=(x, y, 0), half_mm=p.half_mm))
            s = scene.add_volume(Volume(name=f"SiPM{k}", material_id=silicon, parent_volume_id=world,
                                        skin_surface_id=cathode, sensor_id=k,
                                        center_mm=(x, y, hz + 0.5 * p.sipm_thickness_mm),
                                        half_mm=(p.half_mm[0], p.half_mm[1], 0.5 * p.sipm_thickness_mm)))
            if septa:
                scene.add_surface_binding(SurfaceBinding(c, world, wrap))
            crystals.append(c)
            sipms.append(s)
            mats.append(m)
    return CrystalScene(scene, world, crystals, sipms, mats, wrap, cathode, p)


def tapered_crystal_scene(p: CrystalParams | None = None, back_half_xy=(2.0, 2.0), abs_nodes=None,
                          pde_nodes=None) -> CrystalScene:
    """The wrapped crystal tapered from its 3x3 mm front face (z = -10 mm) to a
    back face of half sizes back_half_xy (z = +10 mm): a frustum mesh, its
    four side faces tilted.  The SiPM is a box covering the back face.  The
    crystal's half_mm stays the front face's (the synthetic source box, which
    the frustum contains)."""
    p = p or CrystalParams()
    scene = Scene()
    air, silicon, wrap, cathode = _common_materials_and_surfaces(scene, p, pde_nodes)
    crystal_mat = scene.add_material(_crystal_material(p, abs_nodes))
    world = scene.add_volume(Volume(name="World", material_id=air, center_mm=(0, 0, 0), half_mm=(100, 100, 100)))
    scene.world_volume_id = world
    hz = p.half_mm[2]
    crystal = scene.add_volume(Volume(name="Crystal", material_id=crystal_mat, parent_volume_id=world,
                                      center_mm=(0, 0, 0), half_mm=p.half_mm,
                                      mesh=frustum_mesh(p.half_mm[:2], back_half_xy, hz)))
    sipm = scene.add_volume(Volume(name="SiPM", material_id=silicon, parent_volume_id=world,
                                   skin_surface_id=cathode, sensor_id=0,
                                   center_mm=(0, 0, hz + 0.5 * p.sipm_thickness_mm),
                                   half_mm=(back_half_xy[0], back_half_xy[1], 0.5 * p.sipm_thickness_mm)))
    scene.add_surface_binding(SurfaceBinding(crystal, world, wrap))
    return CrystalScene(scene, world, [crystal], [sipm], [crystal_mat], wrap, cathode, p)


def rotated_crystal_scene(rotation, p: CrystalParams | None = None, abs_nodes=None, pde_nodes=None) -> CrystalScene:
    """crystal_scene with the crystal and its SiPM as box meshes turned by the
    3x3 `rotation` about the origin (the crystal's centre)."""
    cs = crystal_scene(p, abs_nodes, pde_nodes)
    r = np.asarray(rotation, dtype=np.float64).reshape(3, 3)
    for vid in (cs.crystal_ids[0], cs.sipm_ids[0]):
        v = cs.scene.volumes[vid]
        v.mesh = box_mesh(v.half_mm).transformed(r, r @ np.asarray(v.center_mm, dtype=np.float64))
    return cs


L_GUIDE_POLYGON = ((0.0, 0.0), (20.0, 0.0), (20.0, 5.0), (5.0, 5.0), (5.0, 15.0), (0.0, 15.0))


def l_guide_scene(abs_length_mm: float = 30.0):
    """L-shaped light guide of glass (n = 1.5, absorption length abs_length_mm
    in a two-node table) in air: the L polygon L_GUIDE_POLYGON extruded over
    z in [0, 4] mm, a non-convex prism.  Its faces are bare (Fresnel and total
    internal reflection at glass/air, so photons can leave one arm and enter
    the other); a silicon box on the end face x = 20 mm detects what reaches
    it (R = 0, efficiency 0.8).  Photons start uniform in the box
    [6, 19] x [0.5, 4.5] x [0.5, 3.5] mm of the long arm.  Returns (scene,
    source, glass material, detector surface)."""
    scene = Scene()
    air = scene.add_material(Material(name="air", rindex=PropertyTable.constant_value(1.0)))
    glass = scene.add_material(Material(name="glass", rindex=PropertyTable.constant_value(1.5),
                                        abs_length_mm=PropertyTable(np.array([1.5, 4.5], np.float32),
                                                                    np.array([abs_length_mm] * 2, np.float32))))
    si = scene.add_material(Material(name="si", rindex=PropertyTable.constant_value(3.5)))
    detect = scene.add_surface(Surface(name="detect", type=SurfaceType.DielectricMetal,
                                       reflectivity=PropertyTable.constant_value(0.0),
                                       efficiency=PropertyTable.constant_value(0.8)))
    world = scene.add_volume(Volume(name="World", material_id=air, center_mm=(0, 0, 0), half_mm=(60, 60, 60)))
    scene.world_volume_id = world
    guide = scene.add_volume(Volume(name="Guide", material_id=glass, parent_volume_id=world,
                                    mesh=prism_mesh(L_GUIDE_POLYGON, 0.0, 4.0)))
    sensor = scene.add_volume(Volume(name="Sensor", material_id=si, parent_volume_id=world, sensor_id=0,
                                     center_mm=(20.25, 2.5, 2.0), half_mm=(0.25, 2.5, 2.0)))
    scene.add_surface_binding(SurfaceBinding(guide, sensor, detect))
    src = abi.SyntheticSource()
    src.lo_mm[:] = [6.0, 0.5, 0.5]
    src.hi_mm[:] = [19.0, 4.5, 3.5]
    src.wavelength_mean_nm, src.wavelength_sigma_nm, src.decay_time_ns = 420.0, 20.0, 40.0
    src.volume_id = guide
    return scene, src, glass, detect


def ring_scene(n: int = 48, radius_mm: float = 40.0, p: CrystalParams | None = None, wrapped: bool = False):
    """A ring of n crystals (3x3x20 mm box meshes, the long axis radial, centres
    at radius_mm in the z = 0 plane), each turned to face the centre and read
    out by a SiPM (box mesh) on its outer face; bare crystal faces (Fresnel
    with the air between them) unless `wrapped`.  Returns the CrystalScene
    and a synthetic source uniform in a box inside crystal 0 (the one on +x,
    whose faces are axis-aligned)."""
    p = p or CrystalParams()
    scene = Scene()
    air, silicon, wrap, cathode = _common_materials_and_surfaces(scene, p)
    world = scene.add_volume(Volume(name="World", material_id=air, center_mm=(0, 0, 0), half_mm=(100, 100, 100)))
    scene.world_volume_id = world
    hx, hy, hz = p.half_mm
    crystals, sipms, mats = [], [], []
    for k in range(n):
        phi = 2.0 * np.pi * k / n
        radial = np.array([np.cos(phi), np.sin(phi), 0.0])
        tangential = np.array([-np.sin(phi), np.cos(phi), 0.0])
        rot = np.c_[tangential, [0.0, 0.0, 1.0], radial]  # local x, y, z -> tangential, axial, radial
        m = scene.add_material(_crystal_material(p, name=f"crystal{k}"))
        c = scene.add_volume(Volume(name=f"Crystal{k}", material_id=m, parent_volume_id=world, half_mm=p.half_mm,
                                    mesh=box_mesh((hx, hy, hz)).transformed(rot, radius_mm * radial)))
        sp = scene.add_volume(Volume(name=f"SiPM{k}", material_id=silicon, parent_volume_id=world,
                                     skin_surface_id=cathode, sensor_id=k, half_mm=(hx, hy, 0.5 * p.sipm_thickness_mm),
                                     mesh=box_mesh((hx, hy, 0.5 * p.sipm_thickness_mm)).transformed(
                                         rot, (radius_mm + hz + 0.5 * p.sipm_thickness_mm) * radial)))
        if wrapped:
            scene.add_surface_binding(SurfaceBinding(c, world, wrap))
        crystals.append(c)
        sipms.append(sp)
        mats.append(m)
    cs = CrystalScene(scene, world, crystals, sipms, mats, wrap, cathode, p)
    src = abi.SyntheticSource()
    src.lo_mm[:] = [radius_mm - hz + 0.5, -hx + 0.1, -hy + 0.1]
    src.hi_mm[:] = [radius_mm + hz - 0.5, hx - 0.1, hy - 0.1]
    src.wavelength_mean_nm, src.wavelength_sigma_nm, src.decay_time_ns = 420.0, 20.0, 40.0
    src.volume_id = crystals[0]
    return cs, src


def rotation_matrix(axis, angle_rad: float) -> np.ndarray:
    """Rotation by angle_rad about `axis` (Rodrigues)."""
    k = np.asarray(axis, dtype=np.float64)
    k = k / np.linalg.norm(k)
    kx = np.array([[0.0, -k[2], k[1]], [k[2], 0.0, -k[0]], [-k[1], k[0], 0.0]])
    return np.eye(3) + np.sin(angle_rad) * kx + (1.0 - np.cos(angle_rad)) * (kx @ kx)


def point_source(volume_id: int, point_mm, mean_nm: float = 420.0, sigma_nm: float = 20.0,
                 decay_ns: float = 40.0) -> abi.SyntheticSource:
    """Synthetic photons from one point (an empty source box)."""
    src = abi.SyntheticSource()
    src.lo_mm[:] = [float(x) for x in point_mm]
    src.hi_mm[:] = [float(x) for x in point_mm]
    src.wavelength_mean_nm = mean_nm
    src.wavelength_sigma_nm = sigma_nm
    src.decay_time_ns = decay_ns
    src.volume_id = volume_id
    return src


def synthetic_source(cs: CrystalScene, crystal_index: int = 0, mean_nm: float = 420.0, sigma_nm: float = 20.0,
                     decay_ns: float = 40.0) -> abi.SyntheticSource:
    vol = cs.scene.volumes[cs.crystal_ids[crystal_index]]
    lo = [c - h for c, h in zip(vol.center_mm, vol.half_mm)]
    hi = [c + h for c, h in zip(vol.center_mm, vol.half_mm)]
    src = abi.SyntheticSource()
    src.lo_mm[:] = lo
    src.hi_mm[:] = hi
    src.wavelength_mean_nm = mean_nm
    src.wavelength_sigma_nm = sigma_nm
    src.decay_time_ns = decay_ns
    src.volume_id = vol.volume_id
    return src


def lognormal_counts(n_events: int, median: float, sigma: float = 1.5, lo: int = 10, hi: int = 1_000_000,
                     seed: int = 0) -> np.ndarray:
    """clip(round(exp(ln(median) + sigma z)), lo, hi), z ~ N(0, 1), seeded."""
    rng = np.random.default_rng(seed)
    z = rng.standard_normal(n_events)
    return np.clip(np.rint(np.exp(np.log(median) + sigma * z)), lo, hi).astype(np.uint32)


def electron_emissions(cs: CrystalScene, n_events: int, seed: int = 0, energy_mev=1.0,
                       yield_per_kev: float = 30.0, steps: int = 20, range_mm: float = 1.3,
                       first_event_id: int = 0, crystal_index=0):
    """Scintillation emission records of `n_events` electrons entering a crystal on its axis.

    Each electron is a chain of `steps` straight segments from the -z face
    centre with a small random deflection per segment, depositing equal
    energy per segment; photons per segment ~ Poisson(yield * dE).
    `energy_mev` and `crystal_index` are scalars or one value per event.
    Returns (emissions, event_ids, emission_event_indices).
    """
    rng = np.random.default_rng(seed)
    energies = np.broadcast_to(np.asarray(energy_mev, np.float64), (n_events,))
    crystals = np.broadcast_to(np.asarray(crystal_index, np.int64), (n_events,))
    seg = range_mm / steps
    em = np.zeros(n_events * steps, dtype=abi.EMISSION_DTYPE)
    idx = np.repeat(np.arange(n_events, dtype=np.uint32), steps)
    event_ids = np.arange(first_event_id, first_event_id + n_events, dtype=np.uint32)
    beta = 0.94
    v = beta * 299.792458
    for e in range(n_events):
        vol = cs.scene.volumes[cs.crystal_ids[int(crystals[e])]]
        mat = vol.material_id
        start = np.array([vol.center_mm[0], vol.center_mm[1], vol.center_mm[2] - vol.half_mm[2] + 1e-3])
        mean_per_step = yield_per_kev * float(energies[e]) * 1000.0 / steps
        pos = start.copy()
        d = np.array([0.0, 0.0, 1.0])
        first = 0
        counts = rng.poisson(mean_per_step, steps)
        for s in range(steps):
            d = d + rng.normal(0.0, 0.25, 3)
            d /= np.linalg.norm(d)
            if d[2] < 0.2:
                d[2] = 0.2
                d /= np.linalg.norm(d)
            delta = d * seg
            r = em[e * steps + s]
            r["position_mm"] = pos
            r["time_ns"] = s * seg / v
            r["direction"] = d
            r["step_length_mm"] = seg
            r["step_delta_mm"] = delta
            r["pre_velocity_mm_per_ns"] = v
            r["delta_velocity_mm_per_ns"] = -0.01 * v
            r["energy_ev"] = 0.0
            r["pre_mean_photon_count"] = mean_per_step
            r["post_mean_photon_count"] = mean_per_step
            r["decay_time_ns"] = 40.0
            r["rise_time_ns"] = 0.0
            r["charge"] = -1.0
            r["event_id"] = event_ids[e]
            r["first_photon_id"] = first
            r["photon_count"] = counts[s]
            r["volume_id"] = vol.volume_id
            r["material_id"] = mat
            r["spectrum_id"] = 0
            r["type"] = 2
            em[e * steps + s] = r
            first += int(counts[s])
            pos = pos + delta
    return em, event_ids, idx


def array_events(cs: CrystalScene, n_events: int, median: float = 2e4, sigma: float = 1.5, lo: int = 10,
                 hi: int = 1_000_000, seed: int = 0, yield_per_kev: float = 30.0):
    """Uneven electron events in a crystal array: event k deposits the energy of
    lognormal_counts(...)[k] photons (at `yield_per_kev`) in a uniformly chosen
    crystal.  Returns (emissions, event_ids, emission_event_indices, crystal per event)."""
    counts = lognormal_counts(n_events, median, sigma=sigma, lo=lo, hi=hi, seed=seed)
    crystals = np.random.default_rng(seed + 1).integers(0, len(cs.crystal_ids), n_events)
    energies = counts.astype(np.float64) / (yield_per_kev * 1000.0)
    em, ev, idx = electron_emissions(cs, n_events, seed=seed + 2, energy_mev=energies, yield_per_kev=yield_per_kev,
                                     crystal_index=crystals)
    return em, ev, idx, crystals
