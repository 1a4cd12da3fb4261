"""Calibration harness: fit optical parameters of the crystal to a measured
detected count and arrival-time histogram with gradients from the
path-replay adjoint.

Parameters (66): the 32 wavelength-node absorption lengths of the crystal,
its Rayleigh length, the wrap reflectivity and the 32 SiPM efficiency nodes.
The target is produced by transporting with a seed-perturbed parameter set
(each parameter scaled by 1 + 0.2 U(-1, 1)); the fit runs Adam on the
logarithm of the parameters.  The loss is the Pearson chi-square of the
detected count and the 256-bin histogram against the target (Poisson
variances), or with --loss mse the plain sum of squares.  Node-wise absorption
and efficiency trade off against each other at equal wavelength, so the loss
falls while single parameters need not approach the target.

    python -m g4go_amd.calibrate --photons 10000000 --iterations 30
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np
import torch

from . import abi, workloads
from .autograd import chi2_loss, histogram_loss, transport_histogram
from .transport import HipTransportHost, LaunchForm, PhotonTransportConfig


def crystal_layout(cs) -> list:
    m = cs.crystal_materials[0]
    return [(abi.PARAM_ABS_LENGTH, m), (abi.PARAM_RAYLEIGH_LENGTH, m), (abi.PARAM_REFLECTIVITY, cs.wrap_surface),
            (abi.PARAM_EFFICIENCY, cs.cathode_surface)]


def perturbed(theta0: np.ndarray, seed: int, scale: float = 0.2) -> np.ndarray:
    rng = np.random.default_rng(seed)
    theta = theta0 * (1.0 + scale * rng.uniform(-1.0, 1.0, theta0.size))
    return theta


def clamp_probabilities(theta: np.ndarray, n_abs: int, n_pde: int) -> np.ndarray:
    out = theta.copy()
    r = n_abs + 1
    out[r] = min(out[r], 0.999)
    out[r + 1:r + 1 + n_pde] = np.clip(out[r + 1:r + 1 + n_pde], 1e-3, 0.999)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--photons", type=int, default=10_000_000)
    ap.add_argument("--iterations", type=int, default=30)
    ap.add_argument("--lr", type=float, default=0.02)
    ap.add_argument("--loss", choices=("chi2", "mse"), default="chi2")
    ap.add_argument("--beta1", type=float, default=0.5, help="Adam momentum (low: the loss has one stiff direction)")
    ap.add_argument("--decay", type=float, default=0.9, help="learning-rate factor per iteration")
    ap.add_argument("--seed", type=int, default=7)
    args = ap.parse_args(argv)

    cs = workloads.crystal_scene()
    src = workloads.synthetic_source(cs)
    host = HipTransportHost(PhotonTransportConfig(seed=args.seed, sort_hits=False))
    host.set_launch(LaunchForm(persistent=True, block_size=256, chunk=256, refill_threshold=16, blocks_per_cu=4))
    host.prepare_scene(cs.scene)
    n = host.set_parameter_layout(crystal_layout(cs))
    theta0 = host.get_parameters()
    n_abs, n_pde = 32, 32
    theta_star = clamp_probabilities(perturbed(theta0, args.seed + 1), n_abs, n_pde)
    ev = np.array([0], np.uint32)
    counts = np.array([args.photons], np.uint32)
    with torch.no_grad():
        target_host = HipTransportHost(PhotonTransportConfig(seed=args.seed + 1000, sort_hits=False))
        target_host.prepare_scene(cs.scene)
        target_host.set_parameter_layout(crystal_layout(cs))
        n_t, h_t = transport_histogram(torch.from_numpy(theta_star), target_host, cs.scene, src, ev, counts)
        target_host.close()
    log_theta = torch.tensor(np.log(theta0), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([log_theta], lr=args.lr, betas=(args.beta1, 0.999))
    sched = torch.optim.lr_scheduler.ExponentialLR(opt, gamma=args.decay)
    history = []
    for it in range(args.iterations):
        opt.zero_grad()
        theta = torch.exp(log_theta)
        n_det, h = transport_histogram(theta, host, cs.scene, src, ev, counts)
        loss = (chi2_loss if args.loss == "chi2" else histogram_loss)(n_det, h, n_t, h_t)
        loss.backward()
        opt.step()
        sched.step()
        with torch.no_grad():
            clamped = clamp_probabilities(np.exp(log_theta.numpy()), n_abs, n_pde)
            log_theta.copy_(torch.from_numpy(np.log(clamped)))
        err = float(np.median(np.abs(np.exp(log_theta.detach().numpy()) / theta_star - 1.0)))
        history.append({"iteration": it, "loss": float(loss.detach()), "median_rel_param_error": err})
        print(json.dumps(history[-1]), flush=True)
    host.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
