"""Convex mesh solids on the CPU: the oracle's face planes and clipping, and
whole-path closed forms.

The oracle holds a closed convex mesh as the planes of its faces.  What pins
it, independently of its own formulas:
  * a box mesh gives exactly the six axis planes of the box;
  * exit and entry distances and faces agree with a brute-force closest
    triangle hit (Moller-Trumbore in float64) over random convex hulls;
  * solid angles: from a point source inside a tetrahedron with one
    detecting face and absorbing others, the detected fraction is the face's
    solid angle over 4 pi; from a point source outside a detecting
    tetrahedron, it is the solid angle of the faces turned towards the
    point (Van Oosterom-Strackee triangle formula);
  * the tapered crystal (a frustum) with an R = 0 wrap: detected / generated
    = E[solid angle of the back face] / 4 pi * E[PDE] (closed-form rectangle
    solid angle, Gauss-Legendre quadrature over the source box); with R = 1
    every photon ends on the back face: detected / (generated - truncated) =
    E[PDE];
  * the crystal turned by a rotation gives the axis-aligned crystal's
    detected fraction (point source at the same place in the crystal).
The library's own mesh checks (closed, convex, solid) run on the host and are
tested here too.  Counts are binomial; bounds are 4 standard errors.
"""
import numpy as np
import pytest

from g4go_amd import abi, workloads
from g4go_amd.scene import (Material, Mesh, PropertyTable, Scene, Surface, SurfaceBinding, SurfaceType, Volume,
                            box_mesh, prism_mesh)
from oracle import oracle as O

from test_analytic_limits import mean_pde

SEED = 17


# ---------------------------------------------------------------------------
# geometry helpers (float64, independent of the oracle)
def triangle_solid_angle(p, a, b, c):
    """Solid angle of triangle abc seen from p (Van Oosterom and Strackee, 1983)."""
    r1, r2, r3 = a - p, b - p, c - p
    l1, l2, l3 = np.linalg.norm(r1), np.linalg.norm(r2), np.linalg.norm(r3)
    num = abs(np.dot(r1, np.cross(r2, r3)))
    den = l1 * l2 * l3 + np.dot(r1, r2) * l3 + np.dot(r1, r3) * l2 + np.dot(r2, r3) * l1
    return 2.0 * np.arctan2(num, den)


def closest_triangle_hit(verts, tris, p, d, tmin):
    """Closest hit t > tmin over all triangles (Moller-Trumbore), its triangle
    and the smallest barycentric coordinate of the hit (distance to an edge)."""
    v0, v1, v2 = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    e1, e2 = v1 - v0, v2 - v0
    pv = np.cross(d, e2)
    det = np.einsum("ij,ij->i", e1, pv)
    ok = np.abs(det) > 1e-14
    inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
    tv = p - v0
    u = np.einsum("ij,ij->i", tv, pv) * inv
    qv = np.cross(tv, e1)
    v = (qv @ d) * inv
    t = np.einsum("ij,ij->i", e2, qv) * inv
    hit = ok & (u >= 0) & (v >= 0) & (u + v <= 1) & (t > tmin)
    if not hit.any():
        return np.inf, -1, 0.0
    t = np.where(hit, t, np.inf)
    k = int(np.argmin(t))
    return float(t[k]), k, float(min(u[k], v[k], 1 - u[k] - v[k]))


def outward_normal(verts, tri, centroid):
    a, b, c = verts[tri]
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n)
    return n if np.dot(n, a - centroid) > 0 else -n


def random_hull(rng, n_points=40):
    from scipy.spatial import ConvexHull

    pts = rng.normal(size=(n_points, 3))
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    pts = (pts * rng.uniform([3, 5, 8], [4, 7, 12])) @ q.T + rng.uniform(-5, 5, 3)
    hull = ConvexHull(pts)
    return Mesh(pts, hull.simplices.astype(np.uint32))


def mesh_scene(meshes, world_half=200.0):
    """World box of air holding one glass volume per mesh (IDs 1, 2, ...)."""
    sc = Scene()
    air = sc.add_material(Material(name="air", rindex=PropertyTable.constant_value(1.0)))
    glass = sc.add_material(Material(name="glass", rindex=PropertyTable.constant_value(1.5)))
    world = sc.add_volume(Volume(name="world", material_id=air, half_mm=(world_half,) * 3))
    sc.world_volume_id = world
    for k, m in enumerate(meshes):
        sc.add_volume(Volume(name=f"m{k}", material_id=glass, parent_volume_id=world, mesh=m))
    return sc


# ---------------------------------------------------------------------------
def test_box_mesh_gives_exact_axis_planes():
    """A box mesh (also turned by an exact axis permutation) merges to exactly
    the six planes +-e_k . x <= bound, boundary normal = plane normal."""
    half, centre = np.array([1.0, 2.0, 3.0]), np.array([0.5, -4.0, 7.25])
    perm = np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0]])
    for rot in (np.eye(3), perm):
        sc = mesh_scene([box_mesh(half).transformed(rot, centre)])
        for precision in (32, 64):
            planes = O.mesh_planes(sc, 1, precision)
            assert planes.shape == (6, 7)
            got = {(tuple(np.asarray(r[:3], np.float64)), float(r[3])) for r in planes}
            h = np.abs(rot @ half)
            want = set()
            for k in range(3):
                e = np.zeros(3)
                e[k] = 1.0
                want.add((tuple(e), centre[k] + h[k]))
                want.add((tuple(-e + 0.0), -(centre[k] - h[k])))
            assert got == want
            np.testing.assert_array_equal(planes[:, 4:7], planes[:, 0:3])


def test_prism_faces_merge():
    """A hexagonal prism: 20 triangles, 8 faces (coplanar triangles merged)."""
    ang = np.arange(6) * np.pi / 3
    m = prism_mesh(np.c_[5 * np.cos(ang), 5 * np.sin(ang)], -2.0, 3.0)
    assert len(m.indices) == 20
    planes = O.mesh_planes(mesh_scene([m]), 1, 64)
    assert planes.shape == (8, 7)
    np.testing.assert_allclose(np.linalg.norm(planes[:, :3], axis=1), 1.0, atol=1e-15)


@pytest.mark.parametrize("precision", [64, 32])
def test_clipping_matches_closest_triangle_hit(precision):
    """Exit (from inside) and entry (from outside) distances and faces of
    random convex hulls against a brute-force closest triangle hit."""
    rng = np.random.default_rng(5)
    rtol = 1e-9 if precision == 64 else 2e-5
    checked_faces = 0
    for trial in range(6):
        m = random_hull(rng)
        sc = mesh_scene([m])
        planes = O.mesh_planes(sc, 1, 64)
        verts, tris = m.vertices_mm, m.indices.astype(np.int64)
        c = verts.mean(axis=0)
        # exits: points inside (convex combinations of vertices), random directions
        w = rng.dirichlet(np.ones(len(verts)) * 0.3, size=300)
        p_in = w @ verts
        d = rng.normal(size=(300, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        dist, face = O.mesh_trace(sc, 1, p_in, d, entry=False, precision=precision)
        for i in range(300):
            t, k, edge = closest_triangle_hit(verts, tris, p_in[i], d[i], 0.0)
            assert np.isfinite(t)
            assert dist[i] == pytest.approx(t, rel=rtol, abs=rtol * 10), (trial, i)
            if edge > 1e-4:
                np.testing.assert_allclose(planes[face[i], :3], outward_normal(verts, tris[k], c), atol=1e-6)
                checked_faces += 1
        # entries: points outside, aimed at the solid (70%) or anywhere
        p_out = c + rng.normal(size=(300, 3)) * 20.0
        p_out = p_out[[not np.all(planes[:, :3] @ q <= planes[:, 3]) for q in p_out]]
        aim = (rng.dirichlet(np.ones(len(verts)), size=len(p_out)) @ verts) - p_out
        rand = rng.normal(size=aim.shape)
        d = np.where(rng.uniform(size=(len(p_out), 1)) < 0.7, aim, rand)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        dist, face = O.mesh_trace(sc, 1, p_out, d, tmin=1e-4, entry=True, precision=precision)
        hits = 0
        for i in range(len(p_out)):
            t, k, edge = closest_triangle_hit(verts, tris, p_out[i], d[i], 1e-4)
            if edge < 1e-6 and np.isfinite(t):
                continue  # grazing an edge: either answer is right
            assert np.isfinite(dist[i]) == np.isfinite(t), (trial, i, dist[i], t)
            if np.isfinite(t):
                hits += 1
                assert dist[i] == pytest.approx(t, rel=rtol, abs=rtol * 10)
                if edge > 1e-4:
                    np.testing.assert_allclose(planes[face[i], :3], outward_normal(verts, tris[k], c), atol=1e-6)
                    checked_faces += 1
        assert hits > 100
    assert checked_faces > 2000


# ---------------------------------------------------------------------------
TETRA = np.array([[-4.0, -3.0, -2.0], [6.0, -2.0, -1.0], [-1.0, 7.0, 0.0], [0.0, 1.0, 8.0]])
TETRA_TRIS = np.array([[0, 1, 2], [0, 1, 3], [1, 2, 3], [0, 2, 3]])


def tetra_sensor_scene():
    """Glass tetrahedron in air; a silicon prism glued on face (0, 1, 2)
    detects every photon that reaches it, the other faces absorb."""
    sc = Scene()
    air = sc.add_material(Material(name="air", rindex=PropertyTable.constant_value(1.0)))
    glass = sc.add_material(Material(name="glass", rindex=PropertyTable.constant_value(1.5)))
    si = sc.add_material(Material(name="si", rindex=PropertyTable.constant_value(3.5)))
    detect = sc.add_surface(Surface(name="detect", type=SurfaceType.DielectricMetal,
                                    reflectivity=PropertyTable.constant_value(0.0),
                                    efficiency=PropertyTable.constant_value(1.0)))
    absorb = sc.add_surface(Surface(name="absorb", type=SurfaceType.DielectricMetal,
                                    reflectivity=PropertyTable.constant_value(0.0)))
    world = sc.add_volume(Volume(name="world", material_id=air, half_mm=(50.0, 50.0, 50.0)))
    sc.world_volume_id = world
    tet = sc.add_volume(Volume(name="tetra", material_id=glass, parent_volume_id=world,
                               mesh=Mesh(TETRA, TETRA_TRIS)))
    a, b, c = TETRA[0], TETRA[1], TETRA[2]
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n)
    if np.dot(n, TETRA[3] - a) > 0:
        n = -n
    top = np.array([a, b, c]) + 1.0 * n
    verts = np.concatenate([[a, b, c], top])
    tris = [(0, 1, 2), (3, 4, 5)] + [t for i in range(3) for t in
                                     ((i, (i + 1) % 3, 3 + (i + 1) % 3), (i, 3 + (i + 1) % 3, 3 + i))]
    sensor = sc.add_volume(Volume(name="sensor", material_id=si, parent_volume_id=world, sensor_id=0,
                                  mesh=Mesh(verts, np.array(tris))))
    sc.add_surface_binding(SurfaceBinding(tet, sensor, detect))
    sc.add_surface_binding(SurfaceBinding(tet, world, absorb))
    return sc, tet


TETRA_POINT = np.array([0.2, 0.9, 0.6])


def tetra_face_fraction(point=TETRA_POINT):
    return triangle_solid_angle(point, *TETRA[:3]) / (4 * np.pi)


def test_tetrahedron_solid_angles_sum_to_four_pi():
    """The helper: from inside, the four faces cover the sphere."""
    total = sum(triangle_solid_angle(TETRA_POINT, *TETRA[t]) for t in TETRA_TRIS)
    assert total == pytest.approx(4 * np.pi, rel=1e-12)


def binomial_check(k, n, p, what):
    se = np.sqrt(p * (1 - p) / n)
    print(f"{what}: measured {k / n:.6f} expected {p:.6f} ({(k / n - p) / se:+.2f} standard errors)")
    assert abs(k / n - p) < 4 * se


def test_oracle_exit_face_solid_angle():
    sc, tet = tetra_sensor_scene()
    src = workloads.point_source(tet, TETRA_POINT)
    n = 400_000
    r = O.run_synthetic(sc, src, SEED, [0], [n], threads=8)
    st = dict(zip(O.OracleResult.STAT_NAMES, r.stats[0].tolist()))
    assert st["escaped"] == 0 and st["invalid"] == 0 and st["truncated"] == 0
    assert st["detected"] + st["absorbed"] == n
    binomial_check(st["detected"], n, tetra_face_fraction(), "tetrahedron face")


def detecting_tetra_scene():
    """A detecting glass tetrahedron (skin R = 0, efficiency 1) in an air world."""
    sc = Scene()
    air = sc.add_material(Material(name="air", rindex=PropertyTable.constant_value(1.0)))
    glass = sc.add_material(Material(name="glass", rindex=PropertyTable.constant_value(1.5)))
    detect = sc.add_surface(Surface(name="detect", type=SurfaceType.DielectricMetal,
                                    reflectivity=PropertyTable.constant_value(0.0),
                                    efficiency=PropertyTable.constant_value(1.0)))
    world = sc.add_volume(Volume(name="world", material_id=air, half_mm=(50.0, 50.0, 50.0)))
    sc.world_volume_id = world
    sc.add_volume(Volume(name="tetra", material_id=glass, parent_volume_id=world, skin_surface_id=detect,
                         sensor_id=0, mesh=Mesh(TETRA, TETRA_TRIS)))
    return sc, world


OUTSIDE_POINT = np.array([9.0, -7.0, 4.0])


def visible_fraction(point=OUTSIDE_POINT):
    """Solid angle of the faces turned towards the point, over 4 pi."""
    c = TETRA.mean(axis=0)
    total = 0.0
    for t in TETRA_TRIS:
        n = outward_normal(TETRA, t, c)
        if np.dot(n, point - TETRA[t[0]]) > 0:
            total += triangle_solid_angle(point, *TETRA[t])
    return total / (4 * np.pi)


def test_oracle_daughter_entry_solid_angle():
    sc, world = detecting_tetra_scene()
    src = workloads.point_source(world, OUTSIDE_POINT)
    n = 400_000
    r = O.run_synthetic(sc, src, SEED, [0], [n], threads=8)
    st = dict(zip(O.OracleResult.STAT_NAMES, r.stats[0].tolist()))
    assert st["invalid"] == 0 and st["truncated"] == 0 and st["absorbed"] == 0
    assert st["detected"] + st["escaped"] == n
    binomial_check(st["detected"], n, visible_fraction(), "tetrahedron seen from outside")


# ---------------------------------------------------------------------------
def back_face_fraction(source_half=(1.5, 1.5, 10.0), face_half=(2.0, 2.0), order=96):
    """Mean over the source box of (solid angle of the rectangle face_half at
    z = +source_half[2]) / (4 pi)."""
    hx, hy, hz = source_half
    fx, fy = face_half
    g, gw = np.polynomial.legendre.leggauss(order)
    x, y, z = hx * g, hy * g, hz * g
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    W = (hx * gw)[:, None, None] * (hy * gw)[None, :, None] * (hz * gw)[None, None, :]
    d = hz - Z
    omega = np.zeros_like(X)
    for a in (fx - X, fx + X):
        for b in (fy - Y, fy + Y):
            omega += np.arctan(a * b / (d * np.sqrt(a * a + b * b + d * d)))
    return float(np.sum(W * omega / (4 * np.pi)) / np.sum(W))


def tapered_limit_scene(reflectivity):
    cs = workloads.tapered_crystal_scene()
    mat = cs.scene.materials[cs.crystal_materials[0]]
    mat.abs_length_mm = PropertyTable()
    mat.rayleigh_length_mm = PropertyTable()
    cs.scene.surfaces[cs.wrap_surface].reflectivity = PropertyTable.constant_value(reflectivity)
    return cs


def tapered_check(st, reflectivity):
    assert st["escaped"] == 0 and st["invalid"] == 0
    gen = st["generated"]
    assert st["detected"] + st["absorbed"] + st["truncated"] == gen
    if reflectivity == 1.0:
        assert st["truncated"] < 0.02 * gen
        binomial_check(st["detected"], gen - st["truncated"], mean_pde(), "tapered crystal, R = 1")
    else:
        assert st["truncated"] == 0
        binomial_check(st["detected"], gen, back_face_fraction() * mean_pde(), "tapered crystal, R = 0")


def test_back_face_fraction_matches_box_formula():
    """With the face equal to the source box's end face it is the box case."""
    from test_analytic_limits import top_face_fraction

    assert back_face_fraction(face_half=(1.5, 1.5)) == pytest.approx(top_face_fraction(), rel=1e-9)


@pytest.mark.parametrize("reflectivity", [1.0, 0.0])
def test_oracle_tapered_crystal_limits(reflectivity):
    cs = tapered_limit_scene(reflectivity)
    r = O.run_synthetic(cs.scene, workloads.synthetic_source(cs), SEED, [0], [300_000], threads=8)
    tapered_check(dict(zip(O.OracleResult.STAT_NAMES, r.stats[0].tolist())), reflectivity)


# ---------------------------------------------------------------------------
def test_axis_aligned_box_meshes_follow_the_box_paths():
    """The crystal given as unrotated box meshes: same planes as the boxes, so
    nearly every photon ends as it does with box solids (only the last bits
    of positions and reflected directions differ: no snap onto mesh faces,
    reflection by the general formula)."""
    cb = workloads.crystal_scene()
    cm = workloads.rotated_crystal_scene(np.eye(3))
    src = workloads.synthetic_source(cb)
    n = 100_000
    a = O.run_synthetic(cb.scene, src, SEED, [0], [n], threads=8)
    b = O.run_synthetic(cm.scene, src, SEED, [0], [n], threads=8)
    ia, ib = set(a.hits["photon_id"].tolist()), set(b.hits["photon_id"].tolist())
    same = len(ia & ib)
    print(f"detected {len(ia)} vs {len(ib)}, common {same}")
    assert same > 0.99 * max(len(ia), len(ib))


@pytest.mark.parametrize("angle", [0.7, 1.9])
def test_rotated_crystal_matches_axis_aligned(angle):
    """Rotation invariance: a point source at the same place in the crystal."""
    cb = workloads.crystal_scene()
    local = np.array([0.3, -0.2, 1.0])
    rot = workloads.rotation_matrix((1.0, 2.0, 3.0), angle)
    cr = workloads.rotated_crystal_scene(rot)
    n = 300_000
    a = O.run_synthetic(cb.scene, workloads.point_source(cb.crystal_ids[0], local), SEED, [0], [n], threads=8)
    b = O.run_synthetic(cr.scene, workloads.point_source(cr.crystal_ids[0], rot @ local), SEED + 1, [0], [n],
                        threads=8)
    for r in (a, b):
        st = dict(zip(O.OracleResult.STAT_NAMES, r.stats[0].tolist()))
        assert st["escaped"] == 0 and st["invalid"] == 0
    pa, pb = a.stats[0, 0] / n, b.stats[0, 0] / n
    se = np.sqrt(pa * (1 - pa) / n + pb * (1 - pb) / n)
    print(f"axis-aligned {pa:.5f} rotated {pb:.5f} ({(pb - pa) / se:+.2f} standard errors)")
    assert abs(pb - pa) < 4 * se


# ---------------------------------------------------------------------------
def prepare(scene):
    from g4go_amd.transport import HipTransportHost, PhotonTransportConfig

    with HipTransportHost(PhotonTransportConfig(seed=1)) as h:
        h.prepare_scene(scene)


def test_library_accepts_and_rejects_meshes():
    """Host-side scene checks of the library (no device work)."""
    prepare(workloads.tapered_crystal_scene().scene)
    prepare(workloads.rotated_crystal_scene(workloads.rotation_matrix((0, 1, 1), 0.3)).scene)
    prepare(tetra_sensor_scene()[0])
    m = box_mesh((1.0, 1.0, 1.0))
    bad = {
        "degenerate": Mesh(np.r_[m.vertices_mm, m.vertices_mm[:1]], np.r_[m.indices, [[0, 0, 1]]]),
        "index": Mesh(m.vertices_mm, np.r_[m.indices[:-1], [[0, 1, 99]]]),
        "at least 3": Mesh(TETRA[:2], [[0, 1, 1]]),
    }
    for what, mesh in bad.items():
        with pytest.raises(ValueError, match=what):
            prepare(mesh_scene([mesh]))
    # open, flat and single-triangle meshes are open solids (no inside)
    for mesh in (Mesh(m.vertices_mm, m.indices[:-1]), Mesh(np.c_[TETRA[:, :2], np.zeros(4)], TETRA_TRIS),
                 Mesh(TETRA[:3], TETRA_TRIS[:1])):
        sc = mesh_scene([mesh])
        prepare(sc)
        assert O.solid_kind(sc, 1) == 3
    # a dented box (one corner pushed inside) is not convex: a triangle solid
    dented = mesh_scene([Mesh(np.r_[m.vertices_mm[:7], [[0.2, 0.2, 0.2]]], m.indices)])
    prepare(dented)
    assert O.solid_kind(dented, 1) == 2
    normals = np.tile([0.0, 0.0, 0.0], (12, 1))
    with pytest.raises(ValueError, match="normal"):
        prepare(mesh_scene([Mesh(m.vertices_mm, m.indices, normals)]))
    # a convex mesh of more than 1024 faces is held as triangles
    big = mesh_scene([Mesh(*_sphere_mesh(48, 40))])
    prepare(big)
    assert O.solid_kind(big, 1) == 2


def _sphere_mesh(n_lon, n_lat):
    """UV sphere: n_lon * (n_lat - 2) * 2 + 2 * n_lon triangles, all faces distinct."""
    lat = np.linspace(0, np.pi, n_lat + 1)[1:-1]
    lon = np.arange(n_lon) * 2 * np.pi / n_lon
    v = [[0, 0, 10.0], [0, 0, -10.0]]
    for a in lat:
        for b in lon:
            v.append([10 * np.sin(a) * np.cos(b), 10 * np.sin(a) * np.sin(b), 10 * np.cos(a)])
    tri = []
    ring = lambda i, j: 2 + i * n_lon + (j % n_lon)
    for j in range(n_lon):
        tri.append((0, ring(0, j), ring(0, j + 1)))
        tri.append((1, ring(len(lat) - 1, j + 1), ring(len(lat) - 1, j)))
    for i in range(len(lat) - 1):
        for j in range(n_lon):
            tri.append((ring(i, j), ring(i + 1, j), ring(i + 1, j + 1)))
            tri.append((ring(i, j), ring(i + 1, j + 1), ring(i, j + 1)))
    return np.array(v), np.array(tri)


# ---------------------------------------------------------------------------
# Non-convex closed meshes (triangle solids): an L-shaped prism and a hollow
# tube.  Pinned by an analytic inside test, a brute-force closest triangle hit
# and two solid-angle closed forms.
L_POLYGON = [(0.0, 0.0), (20.0, 0.0), (20.0, 5.0), (5.0, 5.0), (5.0, 15.0), (0.0, 15.0)]
TUBE = dict(r_in=3.0, r_out=5.0, half_z=10.0, segments=24)


def l_inside(q):
    x, y, z = q.T
    return (z > 0) & (z < 4) & (((x > 0) & (x < 20) & (y > 0) & (y < 5)) | ((x > 0) & (x < 5) & (y > 0) & (y < 15)))


def tube_inside(q, centre=(0.0, 0.0)):
    """Inside the tube: between the inscribed 24-gons of r_in and r_out, |z| < half_z."""
    n = TUBE["segments"]
    r = np.hypot(q[:, 0] - centre[0], q[:, 1] - centre[1])
    a = np.mod(np.arctan2(q[:, 1] - centre[1], q[:, 0] - centre[0]), 2 * np.pi / n) - np.pi / n
    fac = np.cos(np.pi / n) / np.cos(a)
    return (np.abs(q[:, 2]) < TUBE["half_z"]) & (r < TUBE["r_out"] * fac) & (r > TUBE["r_in"] * fac)


def test_non_convex_meshes_become_triangle_solids():
    from g4go_amd.scene import tube_mesh

    sc = mesh_scene([prism_mesh(L_POLYGON, 0.0, 4.0), tube_mesh(**TUBE), box_mesh((1.0, 1.0, 1.0))])
    assert [O.solid_kind(sc, i) for i in range(4)] == [0, 2, 2, 1]


@pytest.mark.parametrize("precision", [64, 32])
def test_triangle_solid_containment_and_hits(precision):
    """Inside test against the analytic shapes; exit and entry distances
    against a brute-force closest triangle hit (FP64 Moller-Trumbore)."""
    from g4go_amd.scene import tube_mesh

    rng = np.random.default_rng(8)
    lm = prism_mesh(L_POLYGON, 0.0, 4.0)
    tm = tube_mesh(**TUBE).transformed(None, (40.0, 0.0, 0.0))
    sc = mesh_scene([lm, tm])
    q = rng.uniform([-2, -2, -1], [22, 17, 5], size=(20000, 3))
    gap = np.min(np.abs(np.c_[q[:, 2], q[:, 2] - 4, q[:, 0], q[:, 0] - 20, q[:, 0] - 5, q[:, 1], q[:, 1] - 5,
                              q[:, 1] - 15]), axis=1)
    ok = gap > 1e-4
    np.testing.assert_array_equal(O.solid_contains(sc, 1, q, precision)[ok], l_inside(q)[ok])
    q2 = rng.uniform([34, -6, -11], [46, 6, 11], size=(20000, 3))
    np.testing.assert_array_equal(O.solid_contains(sc, 2, q2, precision), tube_inside(q2, (40.0, 0.0)))
    rtol = 1e-9 if precision == 64 else 2e-5
    for vid, mesh, inside in ((1, lm, q[l_inside(q) & ok][:400]), (2, tm, q2[tube_inside(q2, (40.0, 0.0))][:400])):
        d = rng.normal(size=inside.shape)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        dist, face = O.mesh_trace(sc, vid, inside, d, entry=False, precision=precision)
        tris = mesh.indices.astype(np.int64)
        for i in range(len(inside)):
            t, k, edge = closest_triangle_hit(mesh.vertices_mm, tris, inside[i], d[i], 0.0)
            assert dist[i] == pytest.approx(t, rel=rtol, abs=rtol * 10)
            if edge > 1e-4:
                assert face[i] == k  # the triangle, in mesh order
    po = rng.uniform([30, -10, -15], [50, 10, 15], size=(3000, 3))
    po = po[~tube_inside(po, (40.0, 0.0))]
    aim = rng.uniform([35, -5, -10], [45, 5, 10], size=po.shape) - po
    aim /= np.linalg.norm(aim, axis=1, keepdims=True)
    dist, face = O.mesh_trace(sc, 2, po, aim, tmin=1e-4, entry=True, precision=precision)
    hits = 0
    for i in range(len(po)):
        t, k, edge = closest_triangle_hit(tm.vertices_mm, tm.indices.astype(np.int64), po[i], aim[i], 1e-4)
        if edge < 1e-6 and np.isfinite(t):
            continue
        assert np.isfinite(dist[i]) == np.isfinite(t)
        if np.isfinite(t):
            hits += 1
            assert dist[i] == pytest.approx(t, rel=rtol, abs=rtol * 10)
    assert hits > 1000


def l_sensor_scene():
    """Glass L-prism in air; a silicon box on the end face x = 20 of its long
    arm detects every photon reaching it; the other faces absorb."""
    sc = Scene()
    air = sc.add_material(Material(name="air", rindex=PropertyTable.constant_value(1.0)))
    glass = sc.add_material(Material(name="glass", rindex=PropertyTable.constant_value(1.5)))
    si = sc.add_material(Material(name="si", rindex=PropertyTable.constant_value(3.5)))
    detect = sc.add_surface(Surface(name="detect", type=SurfaceType.DielectricMetal,
                                    reflectivity=PropertyTable.constant_value(0.0),
                                    efficiency=PropertyTable.constant_value(1.0)))
    absorb = sc.add_surface(Surface(name="absorb", type=SurfaceType.DielectricMetal,
                                    reflectivity=PropertyTable.constant_value(0.0)))
    world = sc.add_volume(Volume(name="world", material_id=air, half_mm=(60.0, 60.0, 60.0)))
    sc.world_volume_id = world
    lv = sc.add_volume(Volume(name="L", material_id=glass, parent_volume_id=world,
                              mesh=prism_mesh(L_POLYGON, 0.0, 4.0)))
    sensor = sc.add_volume(Volume(name="sensor", material_id=si, parent_volume_id=world, sensor_id=0,
                                  center_mm=(20.25, 2.5, 2.0), half_mm=(0.25, 2.5, 2.0)))
    sc.add_surface_binding(SurfaceBinding(lv, sensor, detect))
    sc.add_surface_binding(SurfaceBinding(lv, world, absorb))
    return sc, lv


L_POINT = np.array([12.0, 2.5, 2.0])


def l_end_face_fraction(p=L_POINT):
    a, b, c, d = (np.array(v) for v in ((20.0, 0.0, 0.0), (20.0, 5.0, 0.0), (20.0, 5.0, 4.0), (20.0, 0.0, 4.0)))
    return (triangle_solid_angle(p, a, b, c) + triangle_solid_angle(p, a, c, d)) / (4 * np.pi)


def test_oracle_l_prism_end_face_solid_angle():
    """From a point of the long arm the whole end face is visible: detected /
    generated = its solid angle / 4 pi."""
    sc, lv = l_sensor_scene()
    n = 400_000
    r = O.run_synthetic(sc, workloads.point_source(lv, L_POINT), SEED, [0], [n], threads=8)
    st = dict(zip(O.OracleResult.STAT_NAMES, r.stats[0].tolist()))
    assert st["escaped"] == 0 and st["invalid"] == 0 and st["detected"] + st["absorbed"] == n
    binomial_check(st["detected"], n, l_end_face_fraction(), "L-prism end face")


def tube_hole_scene():
    """A detecting glass tube (skin R = 0, efficiency 1) in an air world."""
    from g4go_amd.scene import tube_mesh

    sc = Scene()
    air = sc.add_material(Material(name="air", rindex=PropertyTable.constant_value(1.0)))
    glass = sc.add_material(Material(name="glass", rindex=PropertyTable.constant_value(1.5)))
    detect = sc.add_surface(Surface(name="detect", type=SurfaceType.DielectricMetal,
                                    reflectivity=PropertyTable.constant_value(0.0),
                                    efficiency=PropertyTable.constant_value(1.0)))
    world = sc.add_volume(Volume(name="world", material_id=air, half_mm=(60.0, 60.0, 60.0)))
    sc.world_volume_id = world
    sc.add_volume(Volume(name="tube", material_id=glass, parent_volume_id=world, skin_surface_id=detect, sensor_id=0,
                         mesh=tube_mesh(**TUBE)))
    return sc, world


HOLE_POINT = np.array([0.5, 0.3, 2.0])


def tube_wall_fraction(p=HOLE_POINT):
    """From a point in the hole, 1 - (solid angles of the two 24-gon openings) / 4 pi."""
    n, r, hz = TUBE["segments"], TUBE["r_in"], TUBE["half_z"]
    ang = 2 * np.pi * np.arange(n) / n
    omega = 0.0
    for z in (-hz, hz):
        ring = np.c_[r * np.cos(ang), r * np.sin(ang), np.full(n, z)]
        centre = np.array([0.0, 0.0, z])
        omega += sum(triangle_solid_angle(p, centre, ring[k], ring[(k + 1) % n]) for k in range(n))
    return 1.0 - omega / (4 * np.pi)


def test_oracle_tube_hole_wall_fraction():
    """Photons from inside the tube's hole (a point of the world, inside the
    tube's bounding box) enter the non-convex tube unless they leave by an end."""
    sc, world = tube_hole_scene()
    n = 400_000
    r = O.run_synthetic(sc, workloads.point_source(world, HOLE_POINT), SEED, [0], [n], threads=8)
    st = dict(zip(O.OracleResult.STAT_NAMES, r.stats[0].tolist()))
    assert st["invalid"] == 0 and st["absorbed"] == 0 and st["detected"] + st["escaped"] == n
    binomial_check(st["detected"], n, tube_wall_fraction(), "tube wall seen from its hole")


def test_library_triangle_solid_checks():
    """The library accepts the L-prism and the tube, an L-prism with reversed
    triangles (it rewinds them: still a solid), and takes an inconsistently
    wound mesh in two parts and an open tube as open solids."""
    from g4go_amd.scene import tube_mesh

    prepare(l_sensor_scene()[0])
    prepare(tube_hole_scene()[0])
    m = prism_mesh(L_POLYGON, 0.0, 4.0)
    idx = m.indices.copy()
    idx[3] = idx[3][::-1]
    rewound = mesh_scene([Mesh(m.vertices_mm, idx)])
    prepare(rewound)
    assert O.solid_kind(rewound, 1) == 2
    b = box_mesh((1.0, 1.0, 1.0))
    two = Mesh(np.r_[b.vertices_mm, b.vertices_mm + [3.0, 0.0, 0.0]], np.r_[b.indices, b.indices + 8])
    consistent = mesh_scene([two])
    prepare(consistent)  # consistent: two shells are a solid
    assert O.solid_kind(consistent, 1) == 2
    idx = two.indices.copy()
    idx[14] = idx[14][::-1]
    mixed = mesh_scene([Mesh(two.vertices_mm, idx)])
    prepare(mixed)
    assert O.solid_kind(mixed, 1) == 3
    t = tube_mesh(**TUBE)
    opened = mesh_scene([Mesh(t.vertices_mm, t.indices[:-2])])
    prepare(opened)
    assert O.solid_kind(opened, 1) == 3


def test_rewound_mesh_transports_as_consistent_mesh():
    """The L-shaped light guide (Fresnel and total internal reflection on its
    triangles, photons leaving one arm and entering the other) with reversed
    triangles, a scattered third and then all of them, transports photon for
    photon as the consistently wound guide: rewinding restores each
    triangle's outward winding (a wrong one would turn a leaving hit into an
    entering one and change the paths)."""
    scene, src, _, _ = workloads.l_guide_scene()
    ref = O.run_synthetic(scene, src, SEED, [0], [40_000], threads=8)
    assert ref.stats[0][5] == 0 and ref.stats[0][0] > 1000  # no invalid photons, some detected
    rng = np.random.default_rng(5)
    for scattered in (True, False):
        sc2, src2, _, _ = workloads.l_guide_scene()
        m = sc2.volumes[src2.volume_id].mesh
        idx = m.indices.copy()
        sel = rng.random(len(idx)) < 0.33 if scattered else np.ones(len(idx), bool)
        sel[0] = not scattered  # triangle 0 keeps its winding in the scattered case
        idx[sel] = idx[sel][:, [0, 2, 1]]  # reversed with the same first corner, as rewinding restores it
        sc2.volumes[src2.volume_id].mesh = Mesh(m.vertices_mm, idx, m.triangle_normals)
        got = O.run_synthetic(sc2, src2, SEED, [0], [40_000], threads=8)
        assert got.hits.tobytes() == ref.hits.tobytes()
        assert np.array_equal(got.stats, ref.stats)
        prepare(sc2)


def open_cup_scene():
    """A cup: the box mesh |x|, |y|, |z| <= 2 mm without its two z = +2
    triangles (an open solid), filled and surrounded by air; its triangles
    absorb whatever reaches them from inside (R = 0)."""
    sc = Scene()
    air = sc.add_material(Material(name="air", rindex=PropertyTable.constant_value(1.0)))
    absorb = sc.add_surface(Surface(name="absorb", type=SurfaceType.DielectricMetal,
                                    reflectivity=PropertyTable.constant_value(0.0)))
    world = sc.add_volume(Volume(name="world", material_id=air, half_mm=(50.0, 50.0, 50.0)))
    sc.world_volume_id = world
    m = box_mesh((2.0, 2.0, 2.0))
    top = np.all(m.vertices_mm[m.indices][:, :, 2] == 2.0, axis=1)
    assert top.sum() == 2
    cup = sc.add_volume(Volume(name="cup", material_id=air, parent_volume_id=world,
                               mesh=Mesh(m.vertices_mm, m.indices[~top])))
    sc.add_surface_binding(SurfaceBinding(cup, world, absorb))
    return sc, cup


def test_open_cup_escape_fraction():
    """From the centre of the open cup, the photons leaving through its
    opening escape and the rest are absorbed by its walls: the escaped
    fraction is the opening's solid angle over 4 pi, 1/6 (a cube face seen
    from the centre).  Exercises hits from either side, the passage out of an
    open solid with no triangle ahead, and the world's search of an open
    daughter."""
    sc, cup = open_cup_scene()
    assert O.solid_kind(sc, cup) == 3
    prepare(sc)
    n = 300_000
    for precision in (32, 64):
        r = O.run_synthetic(sc, workloads.point_source(cup, (0.0, 0.0, 0.0)), SEED, [0], [n], precision=precision,
                            threads=8)
        st = dict(zip(O.OracleResult.STAT_NAMES, r.stats[0].tolist()))
        assert st["invalid"] == 0 and st["detected"] == 0 and st["absorbed"] + st["escaped"] == n
        binomial_check(st["escaped"], n, 1.0 / 6.0, f"open cup escape (FP{precision})")
