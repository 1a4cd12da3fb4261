"""GPU parity on convex mesh solids: the HIP kernels against the FP32 oracle
on the same Philox streams (hit lists, counters and histograms bit for bit),
and the closed forms of the oracle's own tests at full size on the device.

Scenes: the tapered crystal (frustum mesh, box SiPM on its back face), the
crystal turned by a rotation (crystal and SiPM as box meshes sharing a face),
a tetrahedron with a mesh sensor glued on one face, a detecting tetrahedron
seen from the world (mesh entry), and a world of 40 random convex hulls (a
daughter BVH over mesh solids, face planes read from global memory).
Triangle solids (non-convex meshes): the L-shaped light guide (Fresnel and
TIR on its triangles, photons crossing between its arms through the air),
the L-prism with a sensor on its end face, and the tube seen from its hole.
"""
import sys

import numpy as np
import pytest

from g4go_amd import abi, workloads
from g4go_amd.calibrate import crystal_layout
from g4go_amd.scene import Material, Mesh, PropertyTable, Scene, Surface, SurfaceType, Volume
from g4go_amd.transport import HipTransportHost, LaunchForm, PhotonTransportConfig
from oracle import oracle as O

from test_gpu_parity import assert_same_hits, assert_same_stats
from test_mesh_oracle import (HOLE_POINT, L_POINT, OUTSIDE_POINT, TETRA_POINT, binomial_check, detecting_tetra_scene,
                              l_end_face_fraction, l_sensor_scene, mean_pde, random_hull, tapered_check,
                              tapered_limit_scene, tetra_face_fraction, tetra_sensor_scene, tube_hole_scene,
                              tube_wall_fraction, visible_fraction)

pytestmark = pytest.mark.gpu

PERSISTENT = LaunchForm(persistent=True, block_size=256, chunk=256, refill_threshold=16, blocks_per_cu=4)
PLAIN = LaunchForm(persistent=False)


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def device_run(scene, src, seed, ev, counts, form=PERSISTENT, sort_hits=True):
    with HipTransportHost(PhotonTransportConfig(seed=seed, sort_hits=sort_hits)) as h:
        h.set_launch(form)
        h.enqueue_synthetic(scene, src, np.asarray(ev, np.uint32), np.asarray(counts, np.uint32))
        return h.complete_oldest_batch()


def parity(scene, src, seed, ev, counts, form=PERSISTENT):
    ref = O.run_synthetic(scene, src, seed, ev, counts, threads=8)
    out = device_run(scene, src, seed, ev, counts, form)
    assert out.statistics["invalid"] == 0
    assert_same_hits(out.detections, ref.hits)
    assert_same_stats(out, ref)
    return out


@pytest.mark.parametrize("form", ["plain", "persistent"])
def test_tapered_crystal_bit_exact(form):
    log(f"case: tapered crystal, synthetic, {form}")
    cs = workloads.tapered_crystal_scene()
    out = parity(cs.scene, workloads.synthetic_source(cs), 7, [0, 3], [30000, 4000],
                 PLAIN if form == "plain" else PERSISTENT)
    assert out.statistics["detected"] > 5000


def test_rotated_crystal_bit_exact():
    log("case: rotated crystal (box meshes with a shared face), point source")
    rot = workloads.rotation_matrix((1.0, 2.0, 3.0), 0.7)
    cs = workloads.rotated_crystal_scene(rot)
    src = workloads.point_source(cs.crystal_ids[0], rot @ np.array([0.3, -0.2, 1.0]))
    out = parity(cs.scene, src, 8, [0], [30000])
    assert out.statistics["detected"] > 5000


def test_tetrahedron_face_and_entry_bit_exact():
    log("case: tetrahedron with a mesh sensor; detecting tetrahedron seen from outside")
    sc, tet = tetra_sensor_scene()
    parity(sc, workloads.point_source(tet, TETRA_POINT), 9, [0], [20000])
    sc, world = detecting_tetra_scene()
    parity(sc, workloads.point_source(world, OUTSIDE_POINT), 10, [0], [20000])


def hull_forest(n=40, seed=6):
    """Air world with n random convex glass hulls (every fourth a sensor with a
    partly reflecting skin), absorbing and scattering, on a grid."""
    rng = np.random.default_rng(seed)
    sc = Scene()
    air = sc.add_material(Material(name="air", rindex=PropertyTable.constant_value(1.0)))
    glass = sc.add_material(Material(name="glass", rindex=PropertyTable.constant_value(1.5),
                                     abs_length_mm=PropertyTable.constant_value(40.0),
                                     rayleigh_length_mm=PropertyTable.constant_value(60.0)))
    det = sc.add_surface(Surface(name="det", type=SurfaceType.DielectricMetal,
                                 reflectivity=PropertyTable.constant_value(0.3),
                                 efficiency=PropertyTable.constant_value(0.7)))
    world = sc.add_volume(Volume(name="world", material_id=air, half_mm=(80.0, 80.0, 80.0)))
    sc.world_volume_id = world
    for k in range(n):
        m = random_hull(rng, 24)
        centre = np.array([(k % 4) * 30.0 - 45.0, ((k // 4) % 4) * 30.0 - 45.0, (k // 16) * 30.0 - 30.0])
        m = Mesh(1.2 * (m.vertices_mm - m.vertices_mm.mean(axis=0)) + centre, m.indices)  # radius < 15 mm
        sensor = k % 4 == 0
        sc.add_volume(Volume(name=f"h{k}", material_id=glass, parent_volume_id=world,
                             skin_surface_id=det if sensor else abi.INVALID_ID,
                             sensor_id=k if sensor else abi.INVALID_ID, mesh=m))
    src = abi.SyntheticSource()
    src.lo_mm[:] = [-3.0, -3.0, -18.0]  # a gap between the hulls
    src.hi_mm[:] = [3.0, 3.0, -12.0]
    src.wavelength_mean_nm, src.wavelength_sigma_nm, src.decay_time_ns = 420.0, 20.0, 5.0
    src.volume_id = world
    return sc, src


def test_hull_forest_bvh_bit_exact():
    log("case: 40 random convex hulls, daughter BVH over mesh solids")
    sc, src = hull_forest()
    out = parity(sc, src, 11, [0, 1], [60000, 9000])
    assert out.statistics["detected"] > 800 and out.statistics["escaped"] > 10000


def test_tapered_crystal_emissions_bit_exact():
    log("case: tapered crystal, electron emission records")
    cs = workloads.tapered_crystal_scene()
    em, event_ids, idx = workloads.electron_emissions(cs, 4, seed=3)
    ref = O.run_emissions(cs.scene, em, event_ids, idx, 12)
    with HipTransportHost(PhotonTransportConfig(seed=12, sort_hits=True)) as h:
        h.enqueue_emissions(cs.scene, em, event_ids, idx)
        out = h.complete_oldest_batch()
    assert_same_hits(out.detections, ref.hits)
    assert_same_stats(out, ref)


@pytest.mark.parametrize("reflectivity", [1.0, 0.0])
def test_device_tapered_crystal_limits(reflectivity):
    log(f"case: tapered crystal closed form R={reflectivity}, 2e7 photons")
    cs = tapered_limit_scene(reflectivity)
    out = device_run(cs.scene, workloads.synthetic_source(cs), 33, [0], [20_000_000], sort_hits=False)
    ev = out.event_statistics[0]
    st = {k: int(ev[k]) for k in ("detected", "absorbed", "escaped", "truncated", "invalid")}
    st["generated"] = int(out.statistics["generated"])
    tapered_check(st, reflectivity)


def test_device_tetrahedron_solid_angles():
    log("case: tetrahedron solid angles, 1e7 photons each")
    sc, tet = tetra_sensor_scene()
    n = 10_000_000
    out = device_run(sc, workloads.point_source(tet, TETRA_POINT), 34, [0], [n], sort_hits=False)
    assert out.statistics["invalid"] == 0 and out.statistics["escaped"] == 0
    binomial_check(int(out.statistics["detected"]), n, tetra_face_fraction(), "tetrahedron face")
    sc, world = detecting_tetra_scene()
    out = device_run(sc, workloads.point_source(world, OUTSIDE_POINT), 35, [0], [n], sort_hits=False)
    assert out.statistics["invalid"] == 0 and out.statistics["absorbed"] == 0
    binomial_check(int(out.statistics["detected"]), n, visible_fraction(), "tetrahedron seen from outside")


def test_mesh_adjoint_matches_oracle_replay():
    """The path-replay adjoint on the tapered crystal replays every hit and
    matches the oracle's replay."""
    log("case: adjoint on the tapered crystal vs oracle replay")
    cs = workloads.tapered_crystal_scene()
    src = workloads.synthetic_source(cs)
    ev = np.array([2, 5], np.uint32)
    counts = np.array([40000, 900], np.uint32)
    w = np.random.default_rng(13).normal(size=abi.HIST_SLOTS)
    with HipTransportHost(PhotonTransportConfig(seed=36, sort_hits=True)) as h:
        h.set_launch(PERSISTENT)
        h.prepare_scene(cs.scene)
        n_params = h.set_parameter_layout(crystal_layout(cs))
        h.enqueue_synthetic(cs.scene, src, ev, counts)
        out = h.complete_oldest_batch()
        grad, stats = h.adjoint(w)
    assert stats["replayed"] == len(out.detections) and stats["mismatched"] == 0 and stats["rescored"] == 0
    ref, counters = O.adjoint_synthetic(cs.scene, src, 36, out.detections, w, crystal_layout(cs), n_params)
    assert int(counters[0]) == len(out.detections) and int(counters[1]) == 0
    np.testing.assert_allclose(grad, ref, rtol=1e-9, atol=1e-9 * np.max(np.abs(ref)))


@pytest.mark.parametrize("form", ["plain", "persistent"])
def test_l_guide_triangle_solid_bit_exact(form):
    log(f"case: L-shaped light guide (triangle solid, Fresnel), {form}")
    sc, src, glass, detect = workloads.l_guide_scene()
    out = parity(sc, src, 21, [0, 4], [30000, 3000], PLAIN if form == "plain" else PERSISTENT)
    st = out.statistics
    assert st["detected"] > 1000 and st["escaped"] > 1000


def test_triangle_solid_closed_forms_bit_exact():
    log("case: L-prism end face and tube hole (triangle solids), oracle parity")
    sc, lv = l_sensor_scene()
    parity(sc, workloads.point_source(lv, L_POINT), 22, [0], [20000])
    sc, world = tube_hole_scene()
    parity(sc, workloads.point_source(world, HOLE_POINT), 23, [0], [20000])


def test_device_triangle_solid_solid_angles():
    log("case: triangle-solid solid angles on the device, 1e7 photons each")
    n = 10_000_000
    sc, lv = l_sensor_scene()
    out = device_run(sc, workloads.point_source(lv, L_POINT), 24, [0], [n], sort_hits=False)
    assert out.statistics["invalid"] == 0 and out.statistics["escaped"] == 0
    binomial_check(int(out.statistics["detected"]), n, l_end_face_fraction(), "L-prism end face")
    sc, world = tube_hole_scene()
    out = device_run(sc, workloads.point_source(world, HOLE_POINT), 25, [0], [n], sort_hits=False)
    assert out.statistics["invalid"] == 0 and out.statistics["absorbed"] == 0
    binomial_check(int(out.statistics["detected"]), n, tube_wall_fraction(), "tube wall seen from its hole")


def opened_light_guide():
    """The L-shaped light guide with two of its top face's triangles removed
    (an open solid)."""
    scene, src, glass, detect = workloads.l_guide_scene()
    m = scene.volumes[src.volume_id].mesh
    keep = ~np.all(m.vertices_mm[m.indices][:, :, 2] == 4.0, axis=1)
    keep[np.nonzero(~keep)[0][2:]] = True
    scene.volumes[src.volume_id].mesh = Mesh(m.vertices_mm, m.indices[keep], None)
    return scene, src, glass, detect


@pytest.mark.parametrize("guide", ["closed", "open"])
def test_triangle_solid_adjoint_matches_oracle_replay(guide):
    log(f"case: adjoint on the L-shaped light guide ({guide}) vs oracle replay")
    sc, src, glass, detect = workloads.l_guide_scene() if guide == "closed" else opened_light_guide()
    layout = [(abi.PARAM_ABS_LENGTH, glass), (abi.PARAM_EFFICIENCY, detect)]
    ev = np.array([1, 2], np.uint32)
    counts = np.array([30000, 2000], np.uint32)
    w = np.random.default_rng(14).normal(size=abi.HIST_SLOTS)
    with HipTransportHost(PhotonTransportConfig(seed=37, sort_hits=True)) as h:
        h.set_launch(PERSISTENT)
        h.prepare_scene(sc)
        n_params = h.set_parameter_layout(layout)
        h.enqueue_synthetic(sc, src, ev, counts)
        out = h.complete_oldest_batch()
        grad, stats = h.adjoint(w)
    assert len(out.detections) > 1000
    assert stats["replayed"] == len(out.detections) and stats["mismatched"] == 0 and stats["rescored"] == 0
    ref, counters = O.adjoint_synthetic(sc, src, 37, out.detections, w, layout, n_params)
    assert int(counters[0]) == len(out.detections) and int(counters[1]) == 0
    np.testing.assert_allclose(grad, ref, rtol=1e-9, atol=1e-9 * np.max(np.abs(ref)))


def test_mesh_graph_replays_and_full_size_invariants():
    """Mesh instances under hipGraph replay (several batches, eager work in
    between) and per-event launches equal an eager launch; at 1e8 photons in
    the L-shaped guide every photon ends in one category and hits =
    detections = histogram total."""
    log("case: tapered crystal graph replays and per-event launches; L guide at 1e8 photons")
    cs = workloads.tapered_crystal_scene()
    src = workloads.synthetic_source(cs)
    batches = [workloads.lognormal_counts(n, 300, sigma=1.5, lo=1, hi=20000, seed=s) for n, s in ((30, 1), (20, 2))]
    with HipTransportHost(PhotonTransportConfig(seed=6, sort_hits=True)) as eager, \
            HipTransportHost(PhotonTransportConfig(seed=6, sort_hits=True)) as graphed, \
            HipTransportHost(PhotonTransportConfig(seed=6, sort_hits=True)) as per_event:
        graphed.set_launch(LaunchForm(persistent=True, block_size=256, chunk=256, refill_threshold=16, blocks_per_cu=4,
                                      graph=True))
        per_event.set_launch(LaunchForm(persistent=False, per_event=True))
        for k, counts in enumerate(batches):
            ev = np.arange(100 * k, 100 * k + counts.size, dtype=np.uint32)
            eager.enqueue_synthetic(cs.scene, src, ev, counts)
            want = eager.complete_oldest_batch()
            for h in (graphed, per_event):
                h.enqueue_synthetic(cs.scene, src, ev, counts)
                got = h.complete_oldest_batch()
                assert_same_hits(got.detections, want.detections)
                assert got.event_statistics.tobytes() == want.event_statistics.tobytes()
                np.testing.assert_array_equal(got.histogram, want.histogram)
    sc, gsrc, _, _ = workloads.l_guide_scene()
    rot = workloads.rotation_matrix((1.0, 2.0, 3.0), 0.7)
    rc = workloads.rotated_crystal_scene(rot)
    for name, scene, source in (("L guide", sc, gsrc), ("tapered crystal", cs.scene, src),
                                ("rotated crystal", rc.scene, workloads.point_source(rc.crystal_ids[0], (0, 0, 0)))):
        log(f"  {name}: 1e8 photons")
        out = device_run(scene, source, 38, [0], [100_000_000], sort_hits=False)
        st = out.statistics
        assert st["generated"] == 100_000_000 and st["invalid"] == 0, (name, st["invalid"])
        assert st["detected"] + st["absorbed"] + st["escaped"] + st["truncated"] == st["generated"]
        assert st["n_hits"] == st["detected"] == int(out.histogram.sum())


def nested_mesh_scene():
    """Mesh parents with mesh or box daughters: a glass housing (a box mesh
    turned by a rotation) holding the turned crystal and its SiPM (box meshes,
    wrap and cathode as in the crystal scene), and beside it a glass tube
    (a triangle solid; its border to the world detects, reflects or lets
    through, so photons also cross the hole and re-enter) holding a glass
    bead (a box) in its wall."""
    from g4go_amd.scene import Surface as Surf, SurfaceBinding as Bind, box_mesh, tube_mesh

    rot = workloads.rotation_matrix((1.0, -2.0, 0.5), 1.1)
    cs = workloads.rotated_crystal_scene(rot)
    sc = cs.scene
    glass = sc.add_material(Material(name="glass", rindex=PropertyTable.constant_value(1.5),
                                     abs_length_mm=PropertyTable.constant_value(80.0)))
    housing = sc.add_volume(Volume(name="Housing", material_id=glass, parent_volume_id=cs.world_id,
                                   mesh=box_mesh((3.0, 3.0, 12.0)).transformed(rot, (0.0, 0.0, 0.0))))
    for vid in (cs.crystal_ids[0], cs.sipm_ids[0]):
        sc.volumes[vid].parent_volume_id = housing
    sc.surface_bindings[0] = Bind(cs.crystal_ids[0], housing, cs.wrap_surface)
    det = sc.add_surface(Surf(name="det", type=SurfaceType.DielectricMetal,
                              reflectivity=PropertyTable.constant_value(0.1),
                              transmittance=PropertyTable.constant_value(0.3),
                              efficiency=PropertyTable.constant_value(0.9)))
    tube = sc.add_volume(Volume(name="Tube", material_id=glass, parent_volume_id=cs.world_id, sensor_id=7,
                                mesh=tube_mesh(3.0, 6.0, 10.0, 16).transformed(None, (30.0, 0.0, 0.0))))
    sc.add_surface_binding(Bind(tube, cs.world_id, det))
    sc.add_volume(Volume(name="Bead", material_id=cs.crystal_materials[0], parent_volume_id=tube,
                         center_mm=(34.5, 0.0, 0.0), half_mm=(0.8, 0.8, 3.0)))
    src = abi.SyntheticSource()
    src.lo_mm[:] = [33.6, -0.5, -5.0]   # in the tube wall, around the bead
    src.hi_mm[:] = [35.4, 0.5, 5.0]
    src.wavelength_mean_nm, src.wavelength_sigma_nm, src.decay_time_ns = 420.0, 20.0, 10.0
    src.volume_id = tube
    return cs, src


def test_nested_meshes_bit_exact():
    log("case: nested mesh solids (housing > crystal + SiPM; tube > bead)")
    cs, tube_src = nested_mesh_scene()
    rot_src = workloads.point_source(cs.crystal_ids[0], (0.0, 0.0, 0.0))
    out = parity(cs.scene, rot_src, 39, [0], [20000])
    assert out.statistics["detected"] > 2000
    out = parity(cs.scene, tube_src, 40, [0, 1], [20000, 3000])
    assert out.statistics["detected"] > 1000 and out.statistics["escaped"] > 100


def test_crystal_ring_bit_exact():
    """48 bare crystals turned to face the ring's centre, with SiPMs (96 box
    meshes under the world's daughter BVH): photons leave crystal 0 through
    Fresnel faces and reach its neighbours' sensors."""
    log("case: ring of 48 turned crystals (convex meshes, daughter BVH)")
    cs, src = workloads.ring_scene()
    out = parity(cs.scene, src, 5, [0, 1], [20000, 3000])
    assert out.statistics["detected"] > 5000 and out.statistics["escaped"] > 5000
    assert np.unique(out.detections["sensor_id"]).size > 10


@pytest.mark.parametrize("form", ["plain", "persistent-512"])
def test_crystal_ring_world_source_bit_exact(form):
    """Photons from the air inside the ring: most lanes of a wave search the
    world's 96 mesh daughters at once (each lane testing all of them), later
    only a few (the wave sharing each search); faces whose box meets several
    siblings resolve the neighbour over their sibling lists."""
    log(f"case: ring of 48 turned crystals, source in the world, {form}")
    cs, _ = workloads.ring_scene()
    src = abi.SyntheticSource()
    src.lo_mm[:] = [-20.0, -20.0, -1.0]
    src.hi_mm[:] = [20.0, 20.0, 1.0]
    src.wavelength_mean_nm, src.wavelength_sigma_nm, src.decay_time_ns = 420.0, 20.0, 40.0
    src.volume_id = cs.scene.world_volume_id
    form_ = PLAIN if form == "plain" else LaunchForm(persistent=True, block_size=512, chunk=256, refill_threshold=4,
                                                     blocks_per_cu=2)
    out = parity(cs.scene, src, 21, [0, 2], [60000, 10000], form_)
    assert out.statistics["detected"] > 1000 and out.statistics["escaped"] > 20000
    assert np.unique(out.detections["sensor_id"]).size > 40


def test_large_crystal_ring_bvh_walk_bit_exact():
    """80 crystals and SiPMs (160 daughters, beyond the all-daughter test):
    the world's daughters are searched by the BVH walk."""
    log("case: ring of 80 turned crystals (BVH walk over 160 mesh daughters)")
    cs, src = workloads.ring_scene(n=80, radius_mm=80.0)
    out = parity(cs.scene, src, 23, [0], [20000])
    assert out.statistics["detected"] > 3000 and out.statistics["escaped"] > 3000


def test_rewound_light_guide_bit_exact():
    """The L-shaped light guide with a third of its triangles reversed: the
    library rewinds them as the oracle does (bit for bit against the oracle)."""
    log("case: L-shaped light guide with reversed triangles (rewound on the host)")
    scene, src, _, _ = workloads.l_guide_scene()
    m = scene.volumes[src.volume_id].mesh
    idx = m.indices.copy()
    sel = np.random.default_rng(11).random(len(idx)) < 0.33
    sel[0] = False
    idx[sel] = idx[sel][:, [0, 2, 1]]
    scene.volumes[src.volume_id].mesh = Mesh(m.vertices_mm, idx, m.triangle_normals)
    out = parity(scene, src, 13, [0], [30000])
    assert out.statistics["detected"] > 1000


@pytest.mark.parametrize("form", ["plain", "persistent"])
def test_open_cup_bit_exact(form):
    """The open cup (an open triangle solid): from its centre (absorbed by its
    walls or out through the opening) and from the air outside (entering
    through a wall from the outside, then out again)."""
    from test_mesh_oracle import open_cup_scene

    log(f"case: open cup, {form}")
    sc, cup = open_cup_scene()
    f = PLAIN if form == "plain" else PERSISTENT
    out = parity(sc, workloads.point_source(cup, (0.0, 0.0, 0.0)), 31, [0], [30000], f)
    assert out.statistics["escaped"] > 3000 and out.statistics["absorbed"] > 20000
    out = parity(sc, workloads.point_source(sc.world_volume_id, (3.5, 0.3, 1.0)), 32, [0], [30000], f)
    assert out.statistics["absorbed"] > 1000 and out.statistics["escaped"] > 10000


def test_open_light_guide_bit_exact():
    """The L-shaped light guide with the two triangles of one face removed (an
    open glass solid): Fresnel and total internal reflection on the remaining
    triangles from either side, photons passing out where the face is
    missing."""
    log("case: L-shaped light guide with a face removed (open solid)")
    scene, src, _, _ = opened_light_guide()
    out = parity(scene, src, 33, [0], [30000])
    assert out.statistics["escaped"] > 1000


def test_mixed_box_and_mesh_daughters_bit_exact():
    """A world of 26 daughters, box solids (glass blocks, the SiPM) and the
    tapered crystal (a convex mesh): the all-daughter search tests box and
    mesh daughters alike, shared by the wave (photons from the crystal, few of
    them in the air) or lane by lane (photons from the air)."""
    log("case: tapered crystal among 24 glass blocks (box and mesh daughters)")
    cs = workloads.tapered_crystal_scene()
    sc = cs.scene
    glass = sc.add_material(Material(name="block glass", rindex=PropertyTable.constant_value(1.5),
                                     abs_length_mm=PropertyTable.constant_value(20.0)))
    rng = np.random.default_rng(17)
    blocks = []
    for k in range(24):
        phi = 2.0 * np.pi * k / 24
        centre = (12.0 * np.cos(phi), 12.0 * np.sin(phi), float(rng.uniform(-8.0, 8.0)))
        blocks.append((sc.add_volume(Volume(name=f"block{k}", material_id=glass, parent_volume_id=cs.world_id,
                                            center_mm=centre, half_mm=(1.0, 1.0, 1.5))), centre))
    # photons from a block: most lanes in glass, a few in the air at a time
    out = parity(sc, workloads.point_source(blocks[0][0], blocks[0][1]), 41, [0], [30000])
    assert out.statistics["escaped"] > 10000
    out = parity(sc, workloads.point_source(cs.world_id, (0.0, 6.0, 0.0)), 42, [0], [30000])
    assert out.statistics["escaped"] > 10000 and out.statistics["detected"] > 300
