"""torch.autograd binding of the transport and its path-replay adjoint.

    N, h = transport_histogram(theta, host, scene, source, event_ids, counts)

runs one batch with the bound property tables set to theta (see
HipTransportHost.set_parameter_layout) and returns the detected-photon count
N and the 256-bin arrival-time histogram h as float64 tensors.  Backward
replays the batch's detected photons on the device from their Philox counters
and returns dLoss/dtheta by the likelihood-ratio estimator
(HipTransportHost.adjoint), with per-slot weights dLoss/dN + dLoss/dh_b.
"""
from __future__ import annotations

import numpy as np
import torch

from . import abi


class TransportHistogram(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, host, scene, source, event_ids, counts):
        host.set_parameters(theta.detach().to("cpu", torch.float64).numpy())
        host.enqueue_synthetic(scene, source, event_ids, counts)
        out, _ = host.complete_device()
        ctx.host = host
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
        grad, stats = ctx.host.adjoint(w)
        if stats["mismatched"] != 0:
            raise RuntimeError(f"adjoint replay diverged from the forward batch for {stats['mismatched']} photons")
        return torch.from_numpy(grad).to(ctx.theta_device, ctx.theta_dtype), None, None, None, None, None


def transport_histogram(theta, host, scene, source, event_ids, counts):
    return TransportHistogram.apply(theta, host, scene, source, event_ids, counts)


def histogram_loss(n, h, n_target, h_target):
    """(N - N*)^2 + sum_b (h_b - h*_b)^2."""
    return (n - n_target) ** 2 + torch.sum((h - h_target) ** 2)


def chi2_loss(n, h, n_target, h_target):
    """Pearson chi-square of (N, h) against the target counts (Poisson variances):
    (N - N*)^2 / N* + sum_b (h_b - h*_b)^2 / max(h*_b, 1)."""
    return (n - n_target) ** 2 / torch.clamp(n_target, min=1.0) + torch.sum(
        (h - h_target) ** 2 / torch.clamp(h_target, min=1.0))
