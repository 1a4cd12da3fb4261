"""torch.autograd binding of the transport and its path-replay adjoint.

    N, h = transport_histogram(theta, host, scene, source, event_ids, counts)
    N, h = transport_histogram_emissions(theta, host, scene, emissions, event_ids, emission_event_indices)

run one batch (synthetic photons, or the emission records Geant4 hands over)
with the bound property tables set to theta (see
HipTransportHost.set_parameter_layout) and return the detected-photon count
N and the 256-bin arrival-time histogram h as float64 tensors.  Backward
replays the batch's detected photons on the device from their Philox counters
and returns dLoss/dtheta by the likelihood-ratio estimator
(HipTransportHost.adjoint), with per-slot weights dLoss/dN + dLoss/dh_b.

The adjoint replays the host's last completed batch with the host's current
parameters.  When another batch completed on the host after this forward pass,
or its parameters changed, backward first reruns this batch at this pass's
theta (the Philox streams make the rerun identical) and restores the host's
parameters afterwards.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import abi


@dataclass
class SyntheticBatch:
    source: abi.SyntheticSource
    event_ids: np.ndarray
    counts: np.ndarray

    def enqueue(self, host, scene) -> None:
        host.enqueue_synthetic(scene, self.source, self.event_ids, self.counts)


@dataclass
class EmissionBatch:
    emissions: np.ndarray
    event_ids: np.ndarray
    emission_event_indices: np.ndarray

    def enqueue(self, host, scene) -> None:
        host.enqueue_emissions(scene, self.emissions, self.event_ids, self.emission_event_indices)


def _run(host, scene, batch):
    batch.enqueue(host, scene)
    out, _ = host.complete_device()
    return out


class TransportHistogram(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, host, scene, batch):
        host.set_parameters(theta.detach().to("cpu", torch.float64).numpy())
        out = _run(host, scene, batch)
        ctx.host = host
        ctx.scene = scene
        ctx.batch = batch
        ctx.serial = host.batches_completed
        ctx.theta = host.get_parameters()
        ctx.theta_device = theta.device
        ctx.theta_dtype = theta.dtype
        n = torch.tensor(float(out.statistics["detected"]), dtype=torch.float64)
        h = torch.from_numpy(out.histogram[: abi.HIST_BINS].astype(np.float64))
        return n, h

    @staticmethod
    def backward(ctx, grad_n, grad_h):
        gn = 0.0 if grad_n is None else float(grad_n)
        w = np.full(abi.HIST_SLOTS, gn, np.float64)
        if grad_h is not None:
            w[: abi.HIST_BINS] += grad_h.detach().to("cpu", torch.float64).numpy()
        host = ctx.host
        current = host.get_parameters()
        stale = host.batches_completed != ctx.serial or not np.array_equal(current, ctx.theta)
        if stale:  # replay this pass's batch, not whatever ran since
            host.set_parameters(ctx.theta)
            _run(host, ctx.scene, ctx.batch)
        try:
            grad, stats = host.adjoint(w)
        finally:
            if stale:
                host.set_parameters(current)
        if stats["mismatched"] != 0:
            raise RuntimeError(f"adjoint replay diverged from the forward batch for {stats['mismatched']} photons")
        return torch.from_numpy(grad).to(ctx.theta_device, ctx.theta_dtype), None, None, None


def transport_histogram(theta, host, scene, source, event_ids, counts):
    """(N, h) of a synthetic batch, differentiable in theta."""
    return TransportHistogram.apply(theta, host, scene, SyntheticBatch(source, event_ids, counts))


def transport_histogram_emissions(theta, host, scene, emissions, event_ids, emission_event_indices):
    """(N, h) of a batch of emission records, differentiable in theta."""
    return TransportHistogram.apply(theta, host, scene, EmissionBatch(emissions, event_ids, emission_event_indices))


def histogram_loss(n, h, n_target, h_target):
    """(N - N*)^2 + sum_b (h_b - h*_b)^2."""
    return (n - n_target) ** 2 + torch.sum((h - h_target) ** 2)


def chi2_loss(n, h, n_target, h_target):
    """Pearson chi-square of (N, h) against the target counts (Poisson variances):
    (N - N*)^2 / N* + sum_b (h_b - h*_b)^2 / max(h*_b, 1)."""
    return (n - n_target) ** 2 / torch.clamp(n_target, min=1.0) + torch.sum(
        (h - h_target) ** 2 / torch.clamp(h_target, min=1.0))
