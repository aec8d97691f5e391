"""Attenuation-coefficient calibration with the adjoint of the PE matrix.

Recovers the collimator layers' mu and the bars' mu_total / mu_pe from
point-source projections: each source is one voxel v, its projection is the
column A[:, v].  The loss is the relative squared residual over the chosen
sources,

    L(mu) = sum_{d, v in S} ((A[d][v](mu) - Y[d][v]) / Y[d][v])^2,

so dL/dmu = backward(w) with w = 2 (A - Y) / Y^2 on the sources (0 elsewhere).
The parameters are optimised in log space with Adam.  Everything runs through
the C ABI (smg_pe_forward / smg_pe_backward) on the GPU.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import open_context
from .params import Params


def mu_indices(p: Params) -> tuple[np.ndarray, np.ndarray]:
    """Positions of the calibrated mu in the flat arrays, in gradient order:
    (array id 0 = collimator / 1 = detector, index)."""
    L, N = p.n_layers, p.n_det
    arr = np.concatenate([np.zeros(L, int), np.ones(2 * N, int)])
    idx = np.concatenate([(np.arange(L) + 1) * 10 + 5, 12 * np.arange(N) + 7, 12 * np.arange(N) + 8])
    return arr, idx


def get_mu(p: Params) -> np.ndarray:
    arr, idx = mu_indices(p)
    return np.array([(p.coll if a == 0 else p.det)[i] for a, i in zip(arr, idx)], np.float64)


def set_mu(p: Params, mu: np.ndarray) -> Params:
    q = p.copy()
    arr, idx = mu_indices(q)
    for a, i, m in zip(arr, idx, mu):
        (q.coll if a == 0 else q.det)[i] = np.float32(m)
    return q


@dataclass
class CalibrationResult:
    mu: np.ndarray
    losses: list = field(default_factory=list)
    forward_ms: float = 0.0
    backward_ms: float = 0.0


def calibrate(p_start: Params, target: np.ndarray, sources: np.ndarray, sub=(4, 4, 4),
              steps: int = 200, lr: float = 0.02, device: int = 0) -> CalibrationResult:
    """target: [N][n_sources] projections of the sources (columns of the true
    matrix).  Returns the fitted mu (gradient order) and the loss history."""
    mu = get_mu(p_start)
    theta = np.log(mu)
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)
    b1, b2, eps = 0.9, 0.999, 1e-12
    res = CalibrationResult(mu=mu.copy())
    ctx = open_context(p_start, device)
    ctx.set_pe_subsamples(*sub)
    n_det, n_vox = ctx.dims.n_det, ctx.dims.n_vox
    w = np.zeros((n_det, n_vox), np.float32)
    try:
        for step in range(1, steps + 1):
            p = set_mu(p_start, np.exp(theta))
            ctx.set_geometry(p)
            A, st_f = ctx.pe_forward(0, "fp32")
            pred = A[:, sources].astype(np.float64)
            r = (pred - target) / target
            res.losses.append(float(np.sum(r * r)))
            w[:] = 0.0
            w[:, sources] = (2.0 * r / target).astype(np.float32)
            _, grad, st_b = ctx.pe_backward(w)
            res.forward_ms += st_f["kernel_ms"]
            res.backward_ms += st_b["kernel_ms"]
            g = grad * np.exp(theta)  # d/d(log mu)
            m = b1 * m + (1 - b1) * g
            v = b2 * v + (1 - b2) * g * g
            mh = m / (1 - b1 ** step)
            vh = v / (1 - b2 ** step)
            theta = theta - lr * mh / (np.sqrt(vh) + eps)
    finally:
        ctx.close()
    res.mu = np.exp(theta)
    return res
