"""Pins of the oracle's photon generation and ground-surface sampling against
closed forms (CPU only).

Each distribution the oracle samples is compared with the density that the
mathematics of the reference's algorithm fixes, by a Kolmogorov-Smirnov or
binomial test on 1e5-4e5 samples, and each test carries a negative control
showing that a plausible mistake (a dropped factor, the wrong model, a swapped
fraction) fails it:

* synthetic source: uniform box positions, isotropic directions, polarisation
  perpendicular with a uniform angle, wavelength N(420, 20) nm, time Exp(40 ns);
* scintillation records (OptiXTransportDevice.cu:432-478): the spectrum's
  inverse CDF built from the trapezoid CDF (OptiXTransportHost.cpp:371-410),
  the decay Exp(tau) and the rise-time rejection, whose accepted density is
  (tau_r + tau_d)/tau_d^2 e^{-t/tau_d} (1 - e^{-t/tau_r}) (Geant4
  G4Scintillation's bi-exponential);
* Cerenkov records (:361-430): cos theta = 1/(beta n(E)), the energy density
  proportional to 1 - 1/(beta^2 n(E)^2), the position-fraction density
  proportional to meanPre - f (meanPre - meanPost);
* glisur facets (:492-520): polish 1 gives the normal exactly; otherwise the
  facet is normalize(n + (1 - polish) s) with s uniform in the unit ball, whose
  angle distribution is a ball/cone volume ratio (quadrature);
* unified facets (:522-556): angle density proportional to
  exp(-a^2 / 2 sigma^2) sin a on [0, pi/2];
* unified lobes (:558-580): spike / lobe / backscatter / Lambertian fractions;
* the surfaceRoughness criterion (:843-854): the Lambertian fraction of
  reflections is 1 - exp(-(4 pi sigma n1 cos theta / lambda)^2), and zero
  from the less dense side.
"""
import math

import numpy as np
import pytest
from scipy import integrate, stats

from g4go_amd import abi, workloads
from g4go_amd.scene import (Material, PropertyTable, Scene, Surface, SurfaceBinding, SurfaceFinish, SurfaceModel,
                            SurfaceType, Volume)
from oracle import oracle as O

C_MM_NS = 299.792458
HC_EV_MM = 1.239841984e-3
P_OK = 1e-3     # a correct sampler fails a test with this probability
P_BAD = 1e-8    # a planted mistake must fall below this


def ks_p(x, cdf):
    return stats.kstest(np.asarray(x, np.float64), cdf).pvalue


def tabulated_cdf(pdf, lo, hi, points=20001):
    """CDF of an unnormalised density on [lo, hi] by cumulative Simpson-free trapezoids on a fine grid."""
    g = np.linspace(lo, hi, points)
    f = pdf(g)
    c = np.concatenate([[0.0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(g))])
    c /= c[-1]
    return lambda x: np.interp(x, g, c)


def binomial_ok(k, n, p, sigmas=4.0):
    return abs(k - n * p) <= sigmas * math.sqrt(n * p * (1.0 - p)) + 1.0


# ---------------------------------------------------------------------------
# synthetic source
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("precision", [32, 64])
def test_synthetic_source_distributions(precision):
    cs = workloads.crystal_scene()
    src = workloads.synthetic_source(cs)
    n = 200_000
    g = O.generate_synthetic(cs.scene, src, seed=7, event=3, n=n, precision=precision)
    assert np.all(g[:, 11] == 1)
    lo, hi = np.array(src.lo_mm[:]), np.array(src.hi_mm[:])
    for k in range(3):
        assert ks_p(g[:, k], stats.uniform(lo[k], hi[k] - lo[k]).cdf) > P_OK
    d, p = g[:, 3:6], g[:, 6:9]
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-6)
    assert np.allclose(np.linalg.norm(p, axis=1), 1.0, atol=2e-6)
    assert np.max(np.abs(np.sum(d * p, axis=1))) < 2e-6
    assert ks_p(d[:, 2], stats.uniform(-1.0, 2.0).cdf) > P_OK
    phi = np.arctan2(d[:, 1], d[:, 0])
    assert ks_p(phi, stats.uniform(-np.pi, 2 * np.pi).cdf) > P_OK
    # polarisation angle about the direction, measured from e_theta: uniform
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - d[:, 2] ** 2))
    e_theta = np.c_[d[:, 2] * np.cos(phi), d[:, 2] * np.sin(phi), -sin_t]
    e_phi = np.c_[-np.sin(phi), np.cos(phi), np.zeros(n)]
    psi = np.arctan2(np.sum(p * e_phi, axis=1), np.sum(p * e_theta, axis=1))
    assert ks_p(psi, stats.uniform(-np.pi, 2 * np.pi).cdf) > P_OK
    wl = workloads.HC_EV_NM / g[:, 9]
    assert ks_p(wl, stats.norm(420.0, 20.0).cdf) > P_OK
    assert ks_p(wl, stats.norm(420.0, 21.0).cdf) < P_BAD  # a 5% wrong width fails
    assert ks_p(g[:, 10], stats.expon(scale=40.0).cdf) > P_OK
    assert ks_p(g[:, 10], stats.expon(scale=41.0).cdf) < P_BAD


# ---------------------------------------------------------------------------
# emission records
# ---------------------------------------------------------------------------
def emission_scene(spectrum: PropertyTable, rindex: PropertyTable):
    scene = Scene()
    air = scene.add_material(Material(name="air", rindex=PropertyTable.constant_value(1.0)))
    mat = scene.add_material(Material(name="medium", rindex=rindex,
                                      scintillation_spectrum=[spectrum, PropertyTable(), PropertyTable()]))
    world = scene.add_volume(Volume(name="World", material_id=air, half_mm=(100, 100, 100)))
    scene.world_volume_id = world
    body = scene.add_volume(Volume(name="Body", material_id=mat, parent_volume_id=world, half_mm=(50, 50, 50)))
    return scene, body, mat


def record(volume, material, kind, **fields):
    r = np.zeros(1, abi.EMISSION_DTYPE)
    r["position_mm"] = (1.0, -2.0, 3.0)
    r["direction"] = (0.0, 0.0, 1.0)
    r["pre_velocity_mm_per_ns"] = 200.0
    r["event_id"] = 5
    r["photon_count"] = 1
    r["volume_id"] = volume
    r["material_id"] = material
    r["type"] = kind
    for k, v in fields.items():
        r[k] = v
    return r[0]


def triangle_pdf(e, a=2.5, m=2.9, b=3.5):
    return np.where(e < m, (e - a) / (m - a), (b - e) / (b - m)).clip(0.0)


def test_scintillation_spectrum_inverse_cdf_fine_triangle():
    e = np.linspace(2.5, 3.5, 201).astype(np.float32)
    scene, body, mat = emission_scene(PropertyTable(e, triangle_pdf(e.astype(np.float64)).astype(np.float32)),
                                      PropertyTable.constant_value(1.5))
    rec = record(body, mat, 2, charge=0.0)
    g = O.generate_emission(scene, rec, seed=3, n=200_000)
    exact = stats.triang(c=0.4, loc=2.5, scale=1.0)
    assert ks_p(g[:, 9], exact.cdf) > P_OK
    assert ks_p(g[:, 9], stats.triang(c=0.5, loc=2.5, scale=1.0).cdf) < P_BAD  # a mislocated peak fails


def test_scintillation_spectrum_trapezoid_bins():
    """Coarse spectrum: node masses are the trapezoid integrals, the density constant inside a bin
    (the inverse CDF is linear between nodes)."""
    e = np.array([2.0, 2.5, 3.1, 3.5], np.float32)
    v = np.array([0.0, 1.0, 3.0, 0.0], np.float32)
    scene, body, mat = emission_scene(PropertyTable(e, v), PropertyTable.constant_value(1.5))
    g = O.generate_emission(scene, record(body, mat, 2, charge=0.0), seed=9, n=200_000)
    masses = 0.5 * np.diff(e.astype(np.float64)) * (v[1:] + v[:-1])
    node_cdf = np.concatenate([[0.0], np.cumsum(masses) / masses.sum()])
    assert ks_p(g[:, 9], lambda x: np.interp(x, e.astype(np.float64), node_cdf)) > P_OK
    # the left-endpoint (rectangle) rule a slip would give
    wrong = np.concatenate([[0.0], np.cumsum(np.diff(e) * v[:-1]) / np.sum(np.diff(e) * v[:-1])])
    assert ks_p(g[:, 9], lambda x: np.interp(x, e.astype(np.float64), wrong)) < P_BAD


@pytest.mark.parametrize("charge", [0.0, -1.0])
def test_scintillation_direction_position_and_decay(charge):
    e = np.linspace(2.5, 3.5, 11).astype(np.float32)
    scene, body, mat = emission_scene(PropertyTable(e, np.ones_like(e)), PropertyTable.constant_value(1.5))
    rec = record(body, mat, 2, charge=charge, decay_time_ns=25.0, step_length_mm=0.0,
                 step_delta_mm=(0.6, -0.3, 1.2), time_ns=4.0)
    n = 200_000
    g = O.generate_emission(scene, rec, seed=21, n=n)
    d = g[:, 3:6]
    assert ks_p(d[:, 2], stats.uniform(-1.0, 2.0).cdf) > P_OK
    assert np.max(np.abs(np.sum(d * g[:, 6:9], axis=1))) < 2e-6
    frac = (g[:, 0] - 1.0) / 0.6
    if charge == 0.0:
        assert np.allclose(frac, 1.0, atol=1e-5)  # neutral: the end of the step
    else:
        assert ks_p(frac, stats.uniform(0.0, 1.0).cdf) > P_OK
        assert np.allclose(g[:, 2] - 3.0, 1.2 * frac, atol=1e-5)
    assert ks_p(g[:, 10] - 4.0, stats.expon(scale=25.0).cdf) > P_OK
    assert ks_p(g[:, 10] - 4.0, stats.expon(scale=26.0).cdf) < P_BAD


def rise_decay_cdf(t, tau_r, tau_d):
    """CDF of (tau_r + tau_d)/tau_d^2 e^{-t/tau_d} (1 - e^{-t/tau_r})."""
    tp = tau_d * tau_r / (tau_d + tau_r)
    return (tau_r + tau_d) / tau_d ** 2 * (tau_d * (1 - np.exp(-t / tau_d)) - tp * (1 - np.exp(-t / tp)))


def test_rise_time_rejection_density():
    e = np.linspace(2.5, 3.5, 11).astype(np.float32)
    scene, body, mat = emission_scene(PropertyTable(e, np.ones_like(e)), PropertyTable.constant_value(1.5))
    tau_r, tau_d = 10.0, 40.0
    rec = record(body, mat, 2, charge=0.0, decay_time_ns=tau_d, rise_time_ns=tau_r, step_length_mm=0.0)
    t = O.generate_emission(scene, rec, seed=4, n=200_000)[:, 10]
    assert abs(rise_decay_cdf(1e4, tau_r, tau_d) - 1.0) < 1e-12
    assert ks_p(t, lambda x: rise_decay_cdf(x, tau_r, tau_d)) > P_OK
    # the convolution model (e^{-t/tau_d} - e^{-t/tau_r}) / (tau_d - tau_r) is a different law
    conv = lambda x: 1 - (tau_d * np.exp(-x / tau_d) - tau_r * np.exp(-x / tau_r)) / (tau_d - tau_r)  # noqa: E731
    assert ks_p(t, conv) < P_BAD
    assert ks_p(t, stats.expon(scale=tau_d).cdf) < P_BAD  # the rise time ignored


def cerenkov_setup(rindex: PropertyTable, beta=0.95, pre_mean=10.0, post_mean=30.0):
    e = np.linspace(2.5, 3.5, 11).astype(np.float32)
    scene, body, mat = emission_scene(PropertyTable(e, np.ones_like(e)), rindex)
    v = beta * C_MM_NS
    axis = np.array([0.3, -0.4, 0.866])
    axis = axis / np.linalg.norm(axis)
    rec = record(body, mat, 1, pre_velocity_mm_per_ns=v / 0.999, delta_velocity_mm_per_ns=-2.0 * (v / 0.999 - v),
                 direction=axis, step_length_mm=0.0, step_delta_mm=(0.0, 0.0, 2.0), pre_mean_photon_count=pre_mean,
                 post_mean_photon_count=post_mean)
    beta32 = float(np.float32((np.float32(v / 0.999) + np.float32(0.5) * np.float32(-2.0 * (v / 0.999 - v)))
                              / np.float32(C_MM_NS)))
    return scene, rec, axis, beta32


def test_cerenkov_cone_angle_at_constant_index():
    scene, rec, axis, beta = cerenkov_setup(PropertyTable(np.array([2.0, 4.0], np.float32),
                                                          np.array([1.5, 1.5], np.float32)))
    g = O.generate_emission(scene, rec, seed=8, n=100_000)
    assert np.all(g[:, 11] == 1)
    d, p = g[:, 3:6], g[:, 6:9]
    cos_t = d @ axis
    assert np.max(np.abs(cos_t - 1.0 / (beta * 1.5))) < 2e-6
    # constant n: every energy is accepted, uniform over the table
    assert ks_p(g[:, 9], stats.uniform(2.0, 2.0).cdf) > P_OK
    # polarisation: unit, perpendicular to the photon and in the plane of the axis and the photon
    assert np.max(np.abs(np.sum(d * p, axis=1))) < 2e-6
    assert np.max(np.abs(np.sum(np.cross(axis, d) * p, axis=1))) < 2e-5
    assert np.all(p @ axis < 0.0)  # e_theta of the cone points away from the axis
    # azimuth about the axis: uniform
    u = np.cross(axis, [0.0, 0.0, 1.0])
    u /= np.linalg.norm(u)
    w = np.cross(axis, u)
    az = np.arctan2(d @ w, d @ u)
    assert ks_p(az, stats.uniform(-np.pi, 2 * np.pi).cdf) > P_OK


def test_cerenkov_energy_and_position_densities():
    e0, e1, n0, n1 = 2.0, 4.0, 1.4, 1.6
    scene, rec, axis, beta = cerenkov_setup(PropertyTable(np.array([e0, e1], np.float32),
                                                          np.array([n0, n1], np.float32)))
    g = O.generate_emission(scene, rec, seed=12, n=200_000)
    energy = g[:, 9]

    def n_of(e):
        return n0 + (n1 - n0) * (e - e0) / (e1 - e0)

    cdf = tabulated_cdf(lambda e: 1.0 - 1.0 / (beta * beta * n_of(e) ** 2), e0, e1)
    assert ks_p(energy, cdf) > P_OK
    assert ks_p(energy, stats.uniform(e0, e1 - e0).cdf) < P_BAD  # the sin^2 rejection dropped
    cos_t = g[:, 3:6] @ axis
    assert np.max(np.abs(cos_t - 1.0 / (beta * n_of(energy)))) < 3e-6
    # position fraction along the step: density 10 + 20 f on [0, 1]
    frac = (g[:, 2] - 3.0) / 2.0
    f_cdf = lambda f: (10.0 * f + 10.0 * f * f) / 20.0  # noqa: E731
    assert ks_p(frac, f_cdf) > P_OK
    assert ks_p(frac, lambda f: (30.0 * f - 10.0 * f * f) / 20.0) < P_BAD  # pre and post swapped


# ---------------------------------------------------------------------------
# ground surfaces
# ---------------------------------------------------------------------------
def surface_scene(*surfaces: Surface):
    scene = Scene()
    air = scene.add_material(Material(name="air", rindex=PropertyTable.constant_value(1.0)))
    world = scene.add_volume(Volume(name="World", material_id=air, half_mm=(10, 10, 10)))
    scene.world_volume_id = world
    ids = [scene.add_surface(s) for s in surfaces]
    return scene, ids


def angle_to(v, n):
    return np.arccos(np.clip(v @ n, -1.0, 1.0))


def glisur_angle_cdf(a):
    """P(angle(n + a s, n) <= theta) for s uniform in the unit ball (a < 1): the volume of the ball
    of radius a about the point n inside the cone of half-angle theta about n, over the ball's
    volume.  In cylindrical coordinates about n (height z, radius r) the ball is r <= rb(z) =
    sqrt(a^2 - (z - 1)^2) and the cone r <= z tan(theta), so the volume is the integral over z of
    pi min(rb(z), z tan(theta))^2; the two radii cross where (1 + t^2) z^2 - 2 z + 1 - a^2 = 0."""
    def cdf(theta):
        t = math.tan(theta)
        disc = 1.0 - (1.0 + t * t) * (1.0 - a * a)
        kinks = [] if disc <= 0.0 else [(1.0 + sg * math.sqrt(disc)) / (1.0 + t * t) for sg in (-1.0, 1.0)]
        area = lambda z: math.pi * min(math.sqrt(max(0.0, a * a - (z - 1.0) ** 2)), z * t) ** 2  # noqa: E731
        vol = integrate.quad(area, 1.0 - a, 1.0 + a, points=[k for k in kinks if 1.0 - a < k < 1.0 + a] or None,
                             epsabs=1e-13, epsrel=1e-11, limit=200)[0]
        return vol / (4.0 / 3.0 * math.pi * a ** 3)

    top = math.asin(a)
    grid = np.linspace(0.0, top, 801)
    values = np.array([0.0] + [cdf(x) for x in grid[1:-1]] + [1.0])
    return lambda x: np.interp(x, grid, values)


def test_glisur_polish_one_is_the_normal():
    scene, (s,) = surface_scene(Surface(type=SurfaceType.DielectricMetal, model=SurfaceModel.Glisur,
                                        finish=SurfaceFinish.Ground, model_value=1.0))
    d = np.array([0.6, 0.0, -0.8])
    out, facet, _ = O.surface_reflect(scene, s, d, [0, 0, 1], 2.9, seed=1, n=1000)
    assert np.array_equal(facet, np.tile(np.float32([0, 0, 1]), (1000, 1)))
    assert np.allclose(out, [0.6, 0.0, 0.8], atol=1e-7)


@pytest.mark.parametrize("precision", [32, 64])
def test_glisur_facet_distribution(precision):
    polish = 0.6
    scene, (s,) = surface_scene(Surface(type=SurfaceType.DielectricMetal, model=SurfaceModel.Glisur,
                                        finish=SurfaceFinish.Ground, model_value=polish))
    n = 200_000
    out, facet, _ = O.surface_reflect(scene, s, [0, 0, -1], [0, 0, 1], 2.9, seed=2, n=n, precision=precision)
    facet = facet.astype(np.float64)
    ang = angle_to(facet, np.array([0.0, 0.0, 1.0]))
    assert ks_p(ang, glisur_angle_cdf(1.0 - polish)) > P_OK
    assert ks_p(ang, glisur_angle_cdf(polish)) < P_BAD  # polish used as the smear radius
    az = np.arctan2(facet[:, 1], facet[:, 0])
    assert ks_p(az, stats.uniform(-np.pi, 2 * np.pi).cdf) > P_OK
    # the reflection is specular about the facet
    d = np.array([0.0, 0.0, -1.0])
    expect = d - 2.0 * (facet @ d)[:, None] * facet
    assert np.max(np.abs(out - expect)) < 5e-6


def test_glisur_facets_face_the_photon_at_grazing_incidence():
    scene, (s,) = surface_scene(Surface(type=SurfaceType.DielectricMetal, model=SurfaceModel.Glisur,
                                        finish=SurfaceFinish.Ground, model_value=0.2))
    d = np.array([math.cos(0.1), 0.0, -math.sin(0.1)])
    _, facet, _ = O.surface_reflect(scene, s, d, [0, 0, 1], 2.9, seed=3, n=50_000)
    assert np.all(facet.astype(np.float64) @ d < 0.0)


def unified_angle_cdf(sigma):
    return tabulated_cdf(lambda a: np.exp(-0.5 * a * a / (sigma * sigma)) * np.sin(a), 0.0, 0.5 * math.pi)


@pytest.mark.parametrize("sigma", [0.1, 0.5])
def test_unified_facet_angle_distribution(sigma):
    scene, (s,) = surface_scene(Surface(type=SurfaceType.DielectricMetal, model=SurfaceModel.Unified,
                                        finish=SurfaceFinish.Ground, model_value=sigma,
                                        specular_lobe=PropertyTable.constant_value(1.0)))
    n = 200_000
    nrm = np.array([0.0, 0.6, 0.8])
    out, facet, branch = O.surface_reflect(scene, s, -nrm, nrm, 2.9, seed=5, n=n)
    assert np.all(branch == 1)
    f = facet.astype(np.float64)
    # After 256 rejected attempts the sampler returns the normal itself
    # (OptiXTransportDevice.cu:543-556): an atom of probability (1 - p)^256,
    # p = (2/pi) int_0^{pi/2} exp(-a^2 / 2 sigma^2) sin a da the acceptance
    # of one attempt (alpha = (pi/2) u, accepted when u' <= exp(..) sin alpha).
    p_acc = 2.0 / math.pi * integrate.quad(lambda a: math.exp(-0.5 * a * a / sigma ** 2) * math.sin(a), 0.0,
                                           0.5 * math.pi)[0]
    at_normal = np.all(facet == np.float32(nrm), axis=1)
    assert binomial_ok(int(at_normal.sum()), n, (1.0 - p_acc) ** 256), (at_normal.mean(), (1.0 - p_acc) ** 256)
    ang = angle_to(f[~at_normal], nrm)
    assert ks_p(ang, unified_angle_cdf(sigma)) > P_OK
    assert ks_p(ang, stats.halfnorm(scale=sigma).cdf) < P_BAD  # the sin(alpha) factor dropped
    # lobe reflections are specular about the facet
    expect = -nrm - 2.0 * (f @ -nrm)[:, None] * f
    assert np.max(np.abs(out - expect)) < 5e-6


def test_unified_lobe_fractions_and_directions():
    spike, lobe, back = 0.2, 0.3, 0.1
    scene, (s,) = surface_scene(Surface(type=SurfaceType.DielectricMetal, model=SurfaceModel.Unified,
                                        finish=SurfaceFinish.Ground, model_value=0.2,
                                        specular_spike=PropertyTable.constant_value(spike),
                                        specular_lobe=PropertyTable.constant_value(lobe),
                                        backscatter=PropertyTable.constant_value(back)))
    n = 200_000
    nrm = np.array([0.0, 0.0, 1.0])
    d = np.array([0.5, 0.0, -math.sqrt(0.75)])
    out, facet, branch = O.surface_reflect(scene, s, d, nrm, 2.9, seed=6, n=n)
    out = out.astype(np.float64)
    counts = np.bincount(branch, minlength=4)
    for k, p in enumerate((spike, lobe, back, 1.0 - spike - lobe - back)):
        assert binomial_ok(counts[k], n, p), (k, counts[k] / n, p)
    # a swapped fraction (spike and lobe) fails
    assert not binomial_ok(counts[0], n, lobe)
    specular = d - 2.0 * (d @ nrm) * nrm
    assert np.max(np.abs(out[branch == 0] - specular)) < 1e-6
    assert np.max(np.abs(out[branch == 2] + d)) < 1e-6
    lam = out[branch == 3]
    assert np.all(lam @ nrm > 0.0)
    assert ks_p((lam @ nrm) ** 2, stats.uniform(0.0, 1.0).cdf) > P_OK  # cosine law: cos^2 uniform
    f = facet[branch == 1].astype(np.float64)
    assert np.max(np.abs(out[branch == 1] - (d - 2.0 * (f @ d)[:, None] * f))) < 5e-6


def test_ground_spike_fallback_is_zero():
    """A ground unified surface without lobe tables reflects Lambertian (spike fallback 0 for ground)."""
    scene, (s,) = surface_scene(Surface(type=SurfaceType.DielectricMetal, model=SurfaceModel.Unified,
                                        finish=SurfaceFinish.Ground, model_value=0.2))
    _, _, branch = O.surface_reflect(scene, s, [0, 0, -1], [0, 0, 1], 2.9, seed=7, n=2000)
    assert np.all(branch == 3)


def roughness_scene(roughness_mm, inner_index=1.82, outer_index=1.0):
    scene = Scene()
    outer = scene.add_material(Material(name="outer", rindex=PropertyTable.constant_value(outer_index)))
    inner = scene.add_material(Material(name="inner", rindex=PropertyTable.constant_value(inner_index)))
    world = scene.add_volume(Volume(name="World", material_id=outer, half_mm=(200, 200, 200)))
    scene.world_volume_id = world
    slab = scene.add_volume(Volume(name="Slab", material_id=inner, parent_volume_id=world, half_mm=(100, 100, 5)))
    surf = scene.add_surface(Surface(type=SurfaceType.DielectricDielectric, model=SurfaceModel.Unified,
                                     finish=SurfaceFinish.Ground, model_value=0.0,
                                     specular_spike=PropertyTable.constant_value(1.0),
                                     surface_roughness=PropertyTable.constant_value(roughness_mm)))
    scene.add_surface_binding(SurfaceBinding(slab, world, surf))
    scene.add_surface_binding(SurfaceBinding(world, slab, surf))
    return scene, world, slab


def fresnel_r_s(n1, n2, cos_i):
    sin_t = n1 / n2 * math.sqrt(1.0 - cos_i * cos_i)
    cos_t = math.sqrt(1.0 - sin_t * sin_t)
    return ((n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t)) ** 2


@pytest.mark.parametrize("precision", [32, 64])
def test_roughness_criterion(precision):
    sigma_r = 2.0e-5  # mm
    energy = 2.95
    theta = math.radians(20.0)
    scene, world, slab = roughness_scene(sigma_r)
    d = [math.sin(theta), 0.0, math.cos(theta)]
    state = [0.0, 0.0, 0.0, *d, 0.0, 1.0, 0.0, energy, 0.0]  # s polarisation
    n = 400_000
    out, status = O.step_from(scene, slab, state, seed=31, n=n, precision=precision)
    assert np.all(status == 0)
    reflected = out[:, 11] == slab
    r_s = fresnel_r_s(1.82, 1.0, math.cos(theta))
    assert binomial_ok(int(reflected.sum()), n, r_s)
    dirs = out[reflected, 3:6]
    specular = np.array([math.sin(theta), 0.0, -math.cos(theta)])
    lambertian = np.max(np.abs(dirs - specular), axis=1) > 1e-5
    expo = 4.0 * math.pi * sigma_r * 1.82 * math.cos(theta) / (HC_EV_MM / energy)
    p_lam = 1.0 - math.exp(-expo * expo)
    m = int(reflected.sum())
    assert binomial_ok(int(lambertian.sum()), m, p_lam), (lambertian.mean(), p_lam)
    assert not binomial_ok(int(lambertian.sum()), m, 1.0 - math.exp(-expo))  # the square dropped
    assert np.all(dirs[lambertian, 2] < 0.0)  # back into the slab
    # from the less dense side the criterion does not apply
    scene2, world2, slab2 = roughness_scene(sigma_r, inner_index=1.0, outer_index=1.82)
    out2, status2 = O.step_from(scene2, slab2, state, seed=32, n=50_000, precision=precision)
    refl2 = out2[:, 11] == slab2
    assert refl2.sum() > 0
    assert np.max(np.abs(out2[refl2, 3:6] - specular)) < 1e-5
