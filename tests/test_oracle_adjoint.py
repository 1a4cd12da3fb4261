"""Closed-form pins of the path-replay likelihood-ratio gradient (CPU).

Slab of thickness D = 10 mm and 2000 mm lateral extent, photons starting on
its bottom face with isotropic directions, no Rayleigh scattering and a
refractive index that makes every face a mirror or a sensor (dielectric_metal
surfaces only):
  * bottom face: polished mirror of reflectivity R,
  * top face: SiPM of efficiency p (R = 0),
  * side faces: absorbing (R = 0).
A photon with direction cosine mu toward the top survives with probability
exp(-D / (mu L)); a downward photon is first reflected with probability R.
With E_n the exponential integrals,
  E[N] = N0/2 p (1 + R) E_2(D/L),
  dE[N]/dL = N0/2 p (1 + R) E_1(D/L) D / L^2,
  dE[N]/dR = N0/2 p E_2(D/L),   dE[N]/dp = N0/2 (1 + R) E_2(D/L).
The replay gradient with loss weight 1 in every slot (L = N) estimates these.
"""
import numpy as np
import pytest
from scipy import special

from g4go_amd import abi
from g4go_amd.scene import (Material, PropertyTable, Scene, Surface, SurfaceBinding, SurfaceFinish, SurfaceModel,
                            SurfaceType, Volume)
from oracle import oracle as O

D, L_ABS, R_BOTTOM, PDE = 10.0, 50.0, 0.6, 0.5


def slab_scene():
    sc = Scene()
    air = sc.add_material(Material(name="air", rindex=PropertyTable.constant_value(1.0)))
    slab_mat = sc.add_material(Material(name="slab", rindex=PropertyTable.constant_value(1.5),
                                        abs_length_mm=PropertyTable(np.float32([2.0, 4.0]),
                                                                    np.float32([L_ABS, L_ABS]))))
    si = sc.add_material(Material(name="si", rindex=PropertyTable.constant_value(3.5)))
    black = sc.add_surface(Surface(name="black", type=SurfaceType.DielectricMetal,
                                   reflectivity=PropertyTable.constant_value(0.0)))
    mirror = sc.add_surface(Surface(name="mirror", type=SurfaceType.DielectricMetal, finish=SurfaceFinish.Polished,
                                    model=SurfaceModel.Unified,
                                    reflectivity=PropertyTable.constant_value(R_BOTTOM)))
    cathode = sc.add_surface(Surface(name="cathode", type=SurfaceType.DielectricMetal,
                                     reflectivity=PropertyTable.constant_value(0.0),
                                     efficiency=PropertyTable(np.float32([2.0, 4.0]), np.float32([PDE, PDE]))))
    world = sc.add_volume(Volume(name="world", material_id=air, half_mm=(3000, 3000, 3000)))
    sc.world_volume_id = world
    slab = sc.add_volume(Volume(name="slab", material_id=slab_mat, parent_volume_id=world, half_mm=(1000, 1000, D / 2)))
    top = sc.add_volume(Volume(name="sipm", material_id=si, parent_volume_id=world, skin_surface_id=cathode,
                               sensor_id=0, center_mm=(0, 0, D / 2 + 0.5), half_mm=(1000, 1000, 0.5)))
    bottom = sc.add_volume(Volume(name="mirror", material_id=si, parent_volume_id=world,
                                  center_mm=(0, 0, -D / 2 - 0.5), half_mm=(1000, 1000, 0.5)))
    sc.add_surface_binding(SurfaceBinding(slab, bottom, mirror))
    sc.add_surface_binding(SurfaceBinding(slab, world, black))
    src = abi.SyntheticSource()
    src.lo_mm[:] = [-1.0, -1.0, -D / 2]
    src.hi_mm[:] = [1.0, 1.0, -D / 2]
    src.wavelength_mean_nm, src.wavelength_sigma_nm, src.decay_time_ns = 420.0, 5.0, 0.0
    src.volume_id = slab
    tables = [(abi.PARAM_ABS_LENGTH, slab_mat), (abi.PARAM_REFLECTIVITY, mirror), (abi.PARAM_EFFICIENCY, cathode)]
    return sc, src, tables


@pytest.mark.parametrize("precision", [32, 64])
def test_replay_gradient_closed_form(precision):
    sc, src, tables = slab_scene()
    n0 = 400_000
    run = O.run_synthetic(sc, src, 17, [0], [n0], precision=precision, threads=8)
    a = D / L_ABS
    e2 = special.expn(2, a)
    e1 = special.expn(1, a)
    expect_n = n0 / 2 * PDE * (1 + R_BOTTOM) * e2
    assert abs(len(run.hits) - expect_n) < 5 * np.sqrt(expect_n)
    grad, counters = O.adjoint_synthetic(sc, src, 17, run.hits, np.ones(abi.HIST_SLOTS), tables, 5,
                                         precision=precision)
    assert counters[0] == len(run.hits) and counters[1] == 0 and counters[2] == 0
    d_abs = grad[0] + grad[1]        # both nodes hold the same length
    d_refl = grad[2]
    d_pde = grad[3] + grad[4]
    np.testing.assert_allclose(d_abs, n0 / 2 * PDE * (1 + R_BOTTOM) * e1 * D / L_ABS**2, rtol=0.03)
    np.testing.assert_allclose(d_refl, n0 / 2 * PDE * e2, rtol=0.03)
    np.testing.assert_allclose(d_pde, n0 / 2 * (1 + R_BOTTOM) * e2, rtol=0.03)
    # the efficiency score is 1/p for every detection: exactly N_det / p
    assert d_pde == pytest.approx(len(run.hits) / PDE, rel=1e-12)


def test_replay_detects_path_mismatch():
    """A recorded hit whose path does not replay (wrong time) is counted, not scored."""
    sc, src, tables = slab_scene()
    run = O.run_synthetic(sc, src, 3, [0], [2000], precision=32)
    hits = run.hits.copy()
    hits["time_ns"][0] += 1.0
    grad, counters = O.adjoint_synthetic(sc, src, 3, hits, np.ones(abi.HIST_SLOTS), tables, 5)
    assert counters[1] == 1 and counters[0] == len(hits) - 1


def test_rayleigh_score_has_zero_mean_when_every_photon_is_detected():
    """Cube of side 20 mm, no absorption, Rayleigh length 10 mm, every face a
    sensor of efficiency 1: N = N0 for every L_R, so dN/dL_R = 0, and the replay
    sum of the scores S/L_R^2 - n_R/L_R over all photons has zero mean.  Its
    standard error is about sqrt(N0 E[n_R]) / L_R with E[n_R] = E[S] / L_R; a
    wrong sign or a dropped segment moves it by ~N0 E[n_R] / L_R, hundreds of
    standard errors."""
    lr = 10.0
    sc = Scene()
    air = sc.add_material(Material(name="air", rindex=PropertyTable.constant_value(1.0)))
    mat = sc.add_material(Material(name="scatterer", rindex=PropertyTable.constant_value(1.5),
                                   rayleigh_length_mm=PropertyTable.constant_value(lr)))
    sensor = sc.add_surface(Surface(name="sensor", type=SurfaceType.DielectricMetal,
                                    reflectivity=PropertyTable.constant_value(0.0),
                                    efficiency=PropertyTable.constant_value(1.0)))
    world = sc.add_volume(Volume(name="world", material_id=air, half_mm=(100, 100, 100)))
    sc.world_volume_id = world
    cube = sc.add_volume(Volume(name="cube", material_id=mat, parent_volume_id=world, sensor_id=0,
                                half_mm=(10, 10, 10)))
    sc.add_surface_binding(SurfaceBinding(cube, world, sensor))
    src = abi.SyntheticSource()
    src.lo_mm[:] = [-10.0, -10.0, -10.0]
    src.hi_mm[:] = [10.0, 10.0, 10.0]
    src.wavelength_mean_nm, src.wavelength_sigma_nm, src.decay_time_ns = 420.0, 5.0, 0.0
    src.volume_id = cube
    n0 = 200_000
    run = O.run_synthetic(sc, src, 23, [0], [n0], threads=8)
    assert len(run.hits) == n0
    grad, counters = O.adjoint_synthetic(sc, src, 23, run.hits, np.ones(abi.HIST_SLOTS),
                                         [(abi.PARAM_RAYLEIGH_LENGTH, mat)], 1)
    assert counters[0] == n0 and counters[1] == 0 and counters[2] == 0
    mean_path = float(np.mean(run.hits["time_ns"])) * 299.792458 / 1.5
    mean_scatters = mean_path / lr
    assert mean_scatters > 0.5          # the score terms are not negligible
    scale = n0 * mean_scatters / lr     # size of either term of the score sum
    se = np.sqrt(n0 * mean_scatters) / lr
    assert abs(grad[0]) < 5 * se, (grad[0], se, scale)
    assert scale > 100 * se


def test_emission_replay_closed_form():
    """Direct emission records on the slab's bottom face, each with a fixed
    direction cosine mu_i toward the top: a photon is detected with
    probability p exp(-D / (mu_i L)), so E[N] = sum_i n_i p exp(-D/(mu_i L)),
    dE[N]/dL = sum_i n_i p exp(-D/(mu_i L)) D / (mu_i L^2) and dE[N]/dp = E[N] / p.
    Several records per event exercise the photon-ID -> record lookup."""
    sc, _, tables = slab_scene()
    slab = 1
    mus = [0.9, 0.6, 0.35, 0.8]
    per_record = 25_000
    recs = []
    for ev in (3, 8, 5):
        first = 0
        for k, mu in enumerate(mus):
            r = np.zeros(1, abi.EMISSION_DTYPE)[0]
            r["position_mm"] = (0.3 * k, -0.2 * k, -D / 2)
            r["direction"] = (np.sqrt(1 - mu * mu), 0.0, mu)
            r["polarization"] = (0.0, 1.0, 0.0)
            r["energy_ev"] = 2.9
            r["event_id"] = ev
            r["first_photon_id"] = first
            r["photon_count"] = per_record
            r["volume_id"] = slab
            r["type"] = 0
            first += per_record
            recs.append(r)
    em = np.array(recs, abi.EMISSION_DTYPE)
    event_ids = np.array([3, 8, 5], np.uint32)
    idx = np.repeat(np.arange(3, dtype=np.uint32), len(mus))
    run = O.run_emissions(sc, em, event_ids, idx, 41, threads=4)
    surv = np.array([np.exp(-D / (mu * L_ABS)) for mu in mus])
    n_exp = 3 * per_record * PDE * surv.sum()
    assert abs(len(run.hits) - n_exp) < 5 * np.sqrt(n_exp)
    grad, counters = O.adjoint_emissions(sc, em, 41, run.hits, np.ones(abi.HIST_SLOTS), tables, 5, threads=4)
    assert counters[0] == len(run.hits) and counters[1] == 0 and counters[2] == 0
    d_abs = 3 * per_record * PDE * float(np.sum(surv * D / (np.array(mus) * L_ABS**2)))
    np.testing.assert_allclose(grad[0] + grad[1], d_abs, rtol=0.02)
    assert grad[3] + grad[4] == pytest.approx(len(run.hits) / PDE, rel=1e-12)
    assert grad[2] == 0.0  # no photon is reflected on the way up
    # a hit whose photon ID lies beyond its event's records is not replayed
    bad = run.hits[:3].copy()
    bad["photon_id"][0] = len(mus) * per_record + 7
    _, c2 = O.adjoint_emissions(sc, em, 41, bad, np.ones(abi.HIST_SLOTS), tables, 5)
    assert c2[1] == 1 and c2[0] == 2


STACK_D = (2.0, 3.0, 2.5, 2.5)      # slab thicknesses along z, mm
STACK_L = (30.0, 50.0, 80.0, 40.0)  # their absorption lengths, mm


@pytest.mark.parametrize("precision", [32, 64])
def test_replay_gradient_across_four_materials(precision):
    """Four touching slabs of equal refractive index (no reflection at the
    internal faces), each its own material with absorption length L_k and
    thickness d_k, photons isotropic on the bottom face, a SiPM of
    efficiency p on top and black elsewhere.  With a = sum_k d_k / L_k,
      E[N] = N0/2 p E_2(a),   dE[N]/dL_k = N0/2 p E_1(a) d_k / L_k^2.
    Every detected path crosses all four materials, more than a replay's two
    path slots hold, so this pins the per-material sums of long paths; a
    length credited to the wrong material moves the four slopes apart."""
    sc = Scene()
    air = sc.add_material(Material(name="air", rindex=PropertyTable.constant_value(1.0)))
    si = sc.add_material(Material(name="si", rindex=PropertyTable.constant_value(3.5)))
    black = sc.add_surface(Surface(name="black", type=SurfaceType.DielectricMetal,
                                   reflectivity=PropertyTable.constant_value(0.0)))
    cathode = sc.add_surface(Surface(name="cathode", type=SurfaceType.DielectricMetal,
                                     reflectivity=PropertyTable.constant_value(0.0),
                                     efficiency=PropertyTable.constant_value(PDE)))
    world = sc.add_volume(Volume(name="world", material_id=air, half_mm=(3000, 3000, 3000)))
    sc.world_volume_id = world
    total = sum(STACK_D)
    z = -total / 2
    mats, slabs = [], []
    for k, (d, ln) in enumerate(zip(STACK_D, STACK_L)):
        m = sc.add_material(Material(name=f"slab{k}", rindex=PropertyTable.constant_value(1.5),
                                     abs_length_mm=PropertyTable(np.float32([2.0, 4.0]), np.float32([ln, ln]))))
        v = sc.add_volume(Volume(name=f"slab{k}", material_id=m, parent_volume_id=world,
                                 center_mm=(0, 0, z + d / 2), half_mm=(1000, 1000, d / 2)))
        sc.add_surface_binding(SurfaceBinding(v, world, black))
        mats.append(m)
        slabs.append(v)
        z += d
    sc.add_volume(Volume(name="sipm", material_id=si, parent_volume_id=world, skin_surface_id=cathode, sensor_id=0,
                         center_mm=(0, 0, total / 2 + 0.5), half_mm=(1000, 1000, 0.5)))
    bottom = sc.add_volume(Volume(name="floor", material_id=si, parent_volume_id=world,
                                  center_mm=(0, 0, -total / 2 - 0.5), half_mm=(1000, 1000, 0.5)))
    sc.add_surface_binding(SurfaceBinding(slabs[0], bottom, black))
    src = abi.SyntheticSource()
    src.lo_mm[:] = [-1.0, -1.0, -total / 2]
    src.hi_mm[:] = [1.0, 1.0, -total / 2]
    src.wavelength_mean_nm, src.wavelength_sigma_nm, src.decay_time_ns = 420.0, 5.0, 0.0
    src.volume_id = slabs[0]
    n0 = 400_000
    run = O.run_synthetic(sc, src, 19, [0], [n0], precision=precision, threads=8)
    a = sum(d / ln for d, ln in zip(STACK_D, STACK_L))
    expect_n = n0 / 2 * PDE * special.expn(2, a)
    assert abs(len(run.hits) - expect_n) < 5 * np.sqrt(expect_n)
    tables = [(abi.PARAM_ABS_LENGTH, m) for m in mats]
    grad, counters = O.adjoint_synthetic(sc, src, 19, run.hits, np.ones(abi.HIST_SLOTS), tables, 8,
                                         precision=precision)
    assert counters[0] == len(run.hits) and counters[1] == 0
    assert counters[2] == len(run.hits)  # every path is longer than two materials
    e1 = special.expn(1, a)
    for k, (d, ln) in enumerate(zip(STACK_D, STACK_L)):
        np.testing.assert_allclose(grad[2 * k] + grad[2 * k + 1], n0 / 2 * PDE * e1 * d / ln**2, rtol=0.03,
                                   err_msg=f"slab {k}")
