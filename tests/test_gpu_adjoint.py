"""GPU path-replay adjoint: against the oracle's replay on the same batch, and
against central finite differences with common random numbers (CRN).

Parameter layout of the wrapped crystal (66 values): its 32 absorption-length
nodes, its Rayleigh length, the wrap reflectivity and the SiPM's 32 efficiency
nodes.  The finite differences rerun the batch with the same seed, event ids
and photon ids (so every photon draws the same Philox words) at theta +- eps v
along grouped directions v: all absorption nodes scaled together, the Rayleigh
length, the wrap reflectivity, all efficiency nodes scaled together.
"""
import sys

import numpy as np
import pytest
import torch

from g4go_amd import abi, workloads
from g4go_amd.autograd import histogram_loss, transport_histogram, transport_histogram_emissions
from g4go_amd.calibrate import clamp_probabilities, crystal_layout, perturbed
from g4go_amd.transport import HipTransportHost, LaunchForm, PhotonTransportConfig
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PERSISTENT = LaunchForm(persistent=True, block_size=256, chunk=1024, refill_threshold=16, blocks_per_cu=4)
N_PARAMS = 66
ABS, RAY, WRAP, PDE = slice(0, 32), 32, 33, slice(34, 66)


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def direction(theta, name):
    """Grouped direction v: tables scaled together (relative), R stepped by 0.05 per unit."""
    v = np.zeros_like(theta)
    if name == "wrap_reflectivity":
        v[WRAP] = 0.05
    else:
        sl = {"abs_length": ABS, "rayleigh_length": RAY, "efficiency": PDE}[name]
        v[sl] = theta[sl]
    return v


def open_host(cs, seed):
    h = HipTransportHost(PhotonTransportConfig(seed=seed, sort_hits=False))
    h.set_launch(PERSISTENT)
    h.prepare_scene(cs.scene)
    assert h.set_parameter_layout(crystal_layout(cs)) == N_PARAMS
    return h


def batched_comparison(h, cs, names, eps, photons, batches, first_event=0):
    """Per-batch adjoint d N/d(theta.v) and CRN central differences over `batches`
    batches of `photons` photons (event ids first_event..): returns
    {name: (adjoint mean, finite-difference mean, standard error of their difference)}
    and the mean forward and adjoint kernel milliseconds."""
    src = workloads.synthetic_source(cs)
    theta = h.get_parameters()
    counts = np.array([photons], np.uint32)
    ad = {n: [] for n in names}
    fd = {n: [] for n in names}
    fwd_ms, adj_ms = [], []
    for b in range(batches):
        ev = np.array([first_event + b], np.uint32)
        out = run(h, cs, src, ev, counts, theta)
        grad, stats = h.adjoint(np.ones(abi.HIST_SLOTS))
        assert stats["mismatched"] == 0 and stats["rescored"] == 0
        assert stats["replayed"] == out.statistics["detected"]
        fwd_ms.append(out.statistics["kernel_ms"])
        adj_ms.append(stats["kernel_ms"])
        for n in names:
            v = direction(theta, n)
            ad[n].append(float(grad @ v))
            n_plus = run(h, cs, src, ev, counts, theta + eps * v).statistics["detected"]
            n_minus = run(h, cs, src, ev, counts, theta - eps * v).statistics["detected"]
            fd[n].append((n_plus - n_minus) / (2 * eps))
    h.set_parameters(theta)
    res = {}
    for n in names:
        a, f = np.array(ad[n]), np.array(fd[n])
        se = float(np.std(a - f, ddof=1) / np.sqrt(batches))
        res[n] = (float(a.mean()), float(f.mean()), se)
        log(f"  {n}: adjoint {res[n][0]:.6g} finite difference {res[n][1]:.6g} ratio {res[n][0] / res[n][1]:.5f} "
            f"(standard error of the difference {se / abs(res[n][1]):.4f})")
    return res, float(np.mean(fwd_ms)), float(np.mean(adj_ms))


def run(h, cs, src, ev, counts, theta=None):
    if theta is not None:
        h.set_parameters(theta)
    h.enqueue_synthetic(cs.scene, src, ev, counts)
    out, _ = h.complete_device()
    assert out.statistics["dropped_hits"] == 0
    return out


def test_adjoint_matches_oracle_replay(crystal):
    """Same batch, same hits: the device replay and the FP32 oracle replay give
    the same gradient (FP64 sums in a different order) and replay every hit."""
    log("case: adjoint vs oracle replay, crystal 3 events")
    src = workloads.synthetic_source(crystal)
    ev = np.array([4, 1, 9], np.uint32)
    counts = np.array([60000, 3000, 17], np.uint32)
    w = np.random.default_rng(11).normal(size=abi.HIST_SLOTS)
    with HipTransportHost(PhotonTransportConfig(seed=31, sort_hits=True)) as h:
        h.set_launch(PERSISTENT)
        h.prepare_scene(crystal.scene)
        assert h.set_parameter_layout(crystal_layout(crystal)) == N_PARAMS
        h.enqueue_synthetic(crystal.scene, src, ev, counts)
        out = h.complete_oldest_batch()
        grad, stats = h.adjoint(w)
    assert stats["replayed"] == len(out.detections) and stats["mismatched"] == 0 and stats["rescored"] == 0
    ref, counters = O.adjoint_synthetic(crystal.scene, src, 31, out.detections, w, crystal_layout(crystal), N_PARAMS)
    assert int(counters[0]) == len(out.detections) and int(counters[1]) == 0
    assert stats["steps"] == int(counters[3])
    np.testing.assert_allclose(grad, ref, rtol=1e-9, atol=1e-9 * np.max(np.abs(ref)))


def test_adjoint_matches_finite_differences(crystal):
    """d N / d theta of the crystal along the absorption, wrap-reflectivity and
    efficiency directions: 32 batches of 1e7 photons, eps = 0.05, within 2%;
    the replay costs at most 3x the forward kernel."""
    log("case: adjoint vs CRN finite differences, crystal, 32 x 1e7 photons")
    with open_host(crystal, 5) as h:
        res, fwd_ms, adj_ms = batched_comparison(h, crystal, ("abs_length", "wrap_reflectivity", "efficiency"),
                                                 0.05, 10_000_000, 32)
    log(f"forward {fwd_ms:.3f} ms, adjoint {adj_ms:.3f} ms per 1e7 photons")
    for name, (ad, fd, se) in res.items():
        assert abs(ad / fd - 1) < 0.02, name
    assert adj_ms <= 3.0 * fwd_ms


def test_rayleigh_gradient_turbid_crystal():
    """Rayleigh length 25 mm (scattering shapes the detected count): 64 batches of
    1e7 photons, eps = 0.05, within 2% of the CRN central difference."""
    log("case: Rayleigh gradient, turbid crystal, 64 x 1e7 photons")
    cs = workloads.crystal_scene(workloads.CrystalParams(rayleigh_length_mm=25.0))
    with open_host(cs, 6) as h:
        res, _, _ = batched_comparison(h, cs, ("rayleigh_length", "abs_length"), 0.05, 10_000_000, 64)
    for name, (ad, fd, se) in res.items():
        assert abs(ad / fd - 1) < 0.02, name


def test_rayleigh_gradient_crystal():
    """Rayleigh length 500 mm: d N / d log L_R is about 0.1% of N per unit, so the
    CRN difference is noisy (about 1% standard error at 4.8e9 photons, eps = 0.1);
    the bound is 2% or three standard errors of the batch-paired difference."""
    log("case: Rayleigh gradient, crystal, 48 x 1e8 photons")
    cs = workloads.crystal_scene()
    with open_host(cs, 5) as h:
        res, _, _ = batched_comparison(h, cs, ("rayleigh_length",), 0.1, 100_000_000, 48, first_event=1000)
    ad, fd, se = res["rayleigh_length"]
    assert abs(ad - fd) < max(0.02 * abs(fd), 3.0 * se)


def test_autograd_loss_matches_finite_differences(crystal):
    """MSE loss on (N, 256-bin histogram) against a target run at a +-20% perturbed
    theta: torch.autograd through the adjoint vs CRN central differences at 1e8
    photons, eps = 0.05, within 2% (absorption, wrap reflectivity, efficiency)."""
    log("case: autograd MSE loss vs CRN finite differences, 1e8 photons")
    src = workloads.synthetic_source(crystal)
    ev = np.array([0], np.uint32)
    counts = np.array([100_000_000], np.uint32)
    with open_host(crystal, 8) as h, open_host(crystal, 1008) as ht:
        theta0 = h.get_parameters()
        theta_star = clamp_probabilities(perturbed(theta0, 9), 32, 32)
        with torch.no_grad():
            n_t, h_t = transport_histogram(torch.from_numpy(theta_star), ht, crystal.scene, src, ev, counts)
        theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
        n, hist = transport_histogram(theta, h, crystal.scene, src, ev, counts)
        loss = histogram_loss(n, hist, n_t, h_t)
        loss.backward()
        grad = theta.grad.numpy()

        def loss_at(th):
            with torch.no_grad():
                a, b = transport_histogram(torch.from_numpy(th), h, crystal.scene, src, ev, counts)
                return float(histogram_loss(a, b, n_t, h_t))

        for name in ("abs_length", "wrap_reflectivity", "efficiency"):
            v = direction(theta0, name)
            fd = (loss_at(theta0 + 0.05 * v) - loss_at(theta0 - 0.05 * v)) / 0.1
            ad = float(grad @ v)
            log(f"  {name}: adjoint {ad:.6g} finite difference {fd:.6g} ratio {ad / fd:.5f}")
            assert abs(ad / fd - 1) < 0.02, name


def test_emission_adjoint_finite_differences(crystal):
    """Scintillation emission records (2000 electrons, ~6e7 photons): d N / d theta vs CRN differences."""
    log("case: emission-source adjoint vs CRN finite differences")
    em, event_ids, idx = workloads.electron_emissions(crystal, 2000, seed=21)
    with open_host(crystal, 13) as h:
        theta = h.get_parameters()

        def detected(th):
            h.set_parameters(th)
            h.enqueue_emissions(crystal.scene, em, event_ids, idx)
            out, _ = h.complete_device()
            assert out.statistics["dropped_hits"] == 0
            return out.statistics["detected"]

        n0 = detected(theta)
        grad, stats = h.adjoint(np.ones(abi.HIST_SLOTS))
        assert stats["mismatched"] == 0 and stats["rescored"] == 0 and stats["replayed"] == n0
        for name in ("abs_length", "wrap_reflectivity", "efficiency"):
            v = direction(theta, name)
            fd = (detected(theta + 0.05 * v) - detected(theta - 0.05 * v)) / 0.1
            ad = float(grad @ v)
            log(f"  {name}: adjoint {ad:.6g} finite difference {fd:.6g} ratio {ad / fd:.5f}")
            assert abs(ad / fd - 1) < 0.02, name


def test_autograd_backward_after_another_batch(crystal):
    """Backward replays its own forward batch even when another batch (another
    source, other counts, other parameters) ran on the host in between: the
    gradient equals the one taken right after the forward pass, and the host's
    parameters are left as they were."""
    log("case: autograd backward after an intervening batch")
    src = workloads.synthetic_source(crystal)
    ev, counts = np.array([3], np.uint32), np.array([400_000], np.uint32)
    with open_host(crystal, 17) as h:
        theta0 = h.get_parameters()
        theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
        n, hist = transport_histogram(theta, h, crystal.scene, src, ev, counts)
        (n + hist.sum()).backward()
        direct = theta.grad.numpy().copy()
        theta.grad = None
        n, hist = transport_histogram(theta, h, crystal.scene, src, ev, counts)
        other = theta0.copy()
        other[ABS] *= 0.7
        h.set_parameters(other)
        held = h.get_parameters()
        em, eids, idx = workloads.electron_emissions(crystal, 2, seed=3)
        h.enqueue_emissions(crystal.scene, em, eids, idx)
        h.complete_device()
        (n + hist.sum()).backward()
        np.testing.assert_allclose(theta.grad.numpy(), direct, rtol=1e-9, atol=1e-9 * np.max(np.abs(direct)))
        np.testing.assert_array_equal(h.get_parameters(), held)


def test_autograd_emission_records(crystal):
    """transport_histogram_emissions: torch.autograd over a batch of scintillation
    records gives the host adjoint's gradient of the same batch, and the MSE
    loss gradient agrees with CRN central differences within 2%."""
    log("case: autograd over emission records")
    em, event_ids, idx = workloads.electron_emissions(crystal, 1500, seed=31)
    with open_host(crystal, 19) as h, open_host(crystal, 1019) as ht:
        theta0 = h.get_parameters()
        theta_star = clamp_probabilities(perturbed(theta0, 4), 32, 32)
        with torch.no_grad():
            n_t, h_t = transport_histogram_emissions(torch.from_numpy(theta_star), ht, crystal.scene, em,
                                                     event_ids, idx)
        theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
        n, hist = transport_histogram_emissions(theta, h, crystal.scene, em, event_ids, idx)
        loss = histogram_loss(n, hist, n_t, h_t)
        loss.backward()
        grad = theta.grad.numpy()
        w = np.zeros(abi.HIST_SLOTS)
        w[:] = 2.0 * (float(n.detach()) - float(n_t))
        w[:abi.HIST_BINS] += 2.0 * (hist.detach().numpy() - h_t.numpy())
        direct, stats = h.adjoint(w)
        assert stats["mismatched"] == 0 and stats["rescored"] == 0
        np.testing.assert_allclose(grad, direct, rtol=1e-9, atol=1e-9 * np.max(np.abs(direct)))

        def loss_at(th):
            with torch.no_grad():
                a, b = transport_histogram_emissions(torch.from_numpy(th), h, crystal.scene, em, event_ids, idx)
                return float(histogram_loss(a, b, n_t, h_t))

        for name in ("abs_length", "wrap_reflectivity", "efficiency"):
            v = direction(theta0, name)
            fd = (loss_at(theta0 + 0.05 * v) - loss_at(theta0 - 0.05 * v)) / 0.1
            ad = float(grad @ v)
            log(f"  {name}: adjoint {ad:.6g} finite difference {fd:.6g} ratio {ad / fd:.5f}")
            assert abs(ad / fd - 1) < 0.02, name


def test_parameters_round_trip_and_restore(crystal):
    """set_parameters changes the transport; restoring theta0 restores the batch bit for bit."""
    log("case: parameter round trip")
    src = workloads.synthetic_source(crystal)
    ev = np.array([2], np.uint32)
    counts = np.array([20000], np.uint32)
    with HipTransportHost(PhotonTransportConfig(seed=3, sort_hits=True)) as h:
        h.prepare_scene(crystal.scene)
        h.set_parameter_layout(crystal_layout(crystal))
        theta0 = h.get_parameters()
        a = run(h, crystal, src, ev, counts)
        h.enqueue_synthetic(crystal.scene, src, ev, counts)
        first = h.complete_oldest_batch().detections.copy()
        th = theta0.copy()
        th[PDE] *= 0.5
        h.set_parameters(th)
        h.enqueue_synthetic(crystal.scene, src, ev, counts)
        half = h.complete_oldest_batch().detections.copy()
        assert len(half) < len(first)
        h.set_parameters(theta0)
        h.enqueue_synthetic(crystal.scene, src, ev, counts)
        again = h.complete_oldest_batch().detections.copy()
    assert a.statistics["detected"] == len(first)
    assert first.tobytes() == again.tobytes()


def test_adjoint_matches_fp64_replay(crystal):
    """Config 4 on 1e5 photons: the device gradient against the FP64 CPU replay of
    the same hits.  FP64 paths may leave a few FP32 paths (they are counted as
    mismatches and left out); the rest agree to 1e-3 of the gradient norm."""
    log("case: adjoint vs FP64 oracle replay, 1e5 photons")
    src = workloads.synthetic_source(crystal)
    ev = np.array([0], np.uint32)
    counts = np.array([100_000], np.uint32)
    w = np.random.default_rng(12).normal(size=abi.HIST_SLOTS)
    with open_host(crystal, 41) as h:
        h.enqueue_synthetic(crystal.scene, src, ev, counts)
        out = h.complete_oldest_batch()
        grad, stats = h.adjoint(w)
    ref, counters = O.adjoint_synthetic(crystal.scene, src, 41, out.detections, w, crystal_layout(crystal), N_PARAMS,
                                        precision=64, threads=8)
    log(f"  FP64 replay: {int(counters[0])} replayed, {int(counters[1])} left the FP32 path")
    assert int(counters[1]) <= 0.001 * len(out.detections)
    assert int(counters[0]) + int(counters[1]) == len(out.detections)
    np.testing.assert_allclose(grad, ref, rtol=0, atol=1e-3 * np.linalg.norm(ref))


def test_calibration_reduces_loss():
    """Twelve steps of the calibration harness (66 parameters, 2e6 photons, chi-square
    loss against a +-20% perturbed target) bring the loss down by more than 50x,
    to within a few times the noise level of two independent samples (~2 x 257)."""
    import json
    import subprocess

    log("case: calibration harness, 12 iterations")
    proc = subprocess.run([sys.executable, "-m", "g4go_amd.calibrate", "--photons", "2000000", "--iterations", "12"],
                          capture_output=True, text=True, timeout=110)
    assert proc.returncode == 0, proc.stderr[-2000:]
    hist = [json.loads(l) for l in proc.stdout.splitlines() if l.startswith("{")]
    log(f"  losses {[round(x['loss'], 1) for x in hist]}")
    assert len(hist) == 12 and hist[-1]["loss"] < hist[0]["loss"] / 50
    assert hist[-1]["loss"] < 5 * 2 * 257
