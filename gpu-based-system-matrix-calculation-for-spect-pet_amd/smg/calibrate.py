"""Attenuation-coefficient calibration with the adjoint of the PE matrix.

Recovers the collimator layers' mu and the bars' mu_total / mu_pe from
point-source projections: each source is one voxel v, its projection is the
column A[:, v].  The residuals are relative,

    r[d][v] = (A[d][v](mu) - Y[d][v]) / Y[d][v]  for v in the sources,
    L(mu) = sum r^2,

so dL/dmu = backward(w) with w = 2 r / Y on the sources (0 elsewhere).  The
parameters are theta = log mu.  Everything runs through the C ABI
(smg_pe_forward / smg_pe_backward) on the GPU.

Methods:
  "gauss-newton" (default): damped Gauss-Newton steps
      (J^T J + lam I) delta = -J^T r
  solved by conjugate gradients on the products J^T (J p): J p from the
  tangent (smg_pe_tangent, forward-mode derivative of the matrix along
  mu * p) and J^T u from the adjoint (smg_pe_backward with w = u / Y), two
  passes per product, so the pass count does not depend on the number of
  parameters; a step is taken when it lowers L (lam / 3), else lam * 4 and
  retry.
  "levenberg-marquardt": the same steps with J^T J formed from forward
  differences of the GPU forward (one forward per parameter) and
  lam diag(J^T J) damping, solved directly (small parameter counts).  The problem is ill-conditioned (the
  Jacobian's condition number is ~1.6e4 on the 64-bar geometry: a common
  scale of the bars' mu_pe against the plate mu is the weakest direction), so
  first-order methods crawl along that valley; the Gauss-Newton curvature
  removes it.
  "adam": first-order steps on theta with Adam (lr per step).
  "lbfgs": scipy's L-BFGS-B on theta with the adjoint gradient.
The mu live in float32 arrays (the parameter-file format), so theta is kept
at the log of float32 values and each finite-difference step is divided by
the float32 step actually taken.
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
    n_forward: int = 0
    n_backward: int = 0
    n_tangent: int = 0
    cg_iterations: list = field(default_factory=list)


class _Problem:
    """Forward / adjoint evaluations of the calibration loss on one context."""

    def __init__(self, p_start: Params, target: np.ndarray, sources: np.ndarray, sub, precision,
                 device: int, res: CalibrationResult):
        self.p = p_start
        self.target = np.asarray(target, np.float64)
        self.sources = np.asarray(sources)
        self.prec = precision
        self.res = res
        self.ctx = open_context(p_start, device)
        self.ctx.set_pe_subsamples(*sub)
        self.w = np.zeros((self.ctx.dims.n_det, self.ctx.dims.n_vox), np.float32)

    def close(self):
        self.ctx.close()

    def residuals(self, mu32: np.ndarray) -> np.ndarray:
        self.ctx.set_geometry(set_mu(self.p, mu32))
        A, st = self.ctx.pe_forward(0, self.prec)
        self.res.forward_ms += st["kernel_ms"]
        self.res.n_forward += 1
        return (A[:, self.sources].astype(np.float64) - self.target) / self.target

    def jvp(self, mu32: np.ndarray, p: np.ndarray) -> np.ndarray:
        """J_theta p (relative residuals) at the context's geometry: the
        tangent along dmu = mu * p on the source columns."""
        dA, _, st = self.ctx.pe_tangent(mu32 * p, 0, self.prec)
        self.res.forward_ms += st["kernel_ms"]
        self.res.n_tangent += 1
        return dA[:, self.sources].astype(np.float64) / self.target

    def vjp(self, mu32: np.ndarray, u: np.ndarray) -> np.ndarray:
        """J_theta^T u at the context's geometry: the adjoint with w = u / Y on
        the source columns, times mu."""
        self.w[:] = 0.0
        self.w[:, self.sources] = (u / self.target).astype(np.float32)
        _, g, st = self.ctx.pe_backward(self.w, precision=self.prec)
        self.res.backward_ms += st["kernel_ms"]
        self.res.n_backward += 1
        return g * mu32

    def grad_theta(self, mu32: np.ndarray, r: np.ndarray) -> np.ndarray:
        """dL/dtheta at the mu of the last residuals call (the context's geometry)."""
        self.w[:] = 0.0
        self.w[:, self.sources] = (2.0 * r / self.target).astype(np.float32)
        _, g, st = self.ctx.pe_backward(self.w, precision=self.prec)
        self.res.backward_ms += st["kernel_ms"]
        self.res.n_backward += 1
        return g * mu32


def _f32(mu: np.ndarray) -> np.ndarray:
    return mu.astype(np.float32).astype(np.float64)


def calibrate(p_start: Params, target: np.ndarray, sources: np.ndarray, sub=(4, 4, 4),
              steps: int = 20, lr: float = 0.02, device: int = 0,
              method: str = "gauss-newton", precision: str = "fp32",
              fd_step: float = 1e-4, lam0: float = 0.1, cg_iters: int = 60,
              cg_tol: float = 1e-4) -> CalibrationResult:
    """target: [N][n_sources] projections of the sources (columns of the true
    matrix).  steps: iterations of the method (for "lbfgs", evaluations).
    Returns the fitted mu (gradient order) and the loss after each step."""
    mu = _f32(get_mu(p_start))
    res = CalibrationResult(mu=mu.copy())
    prob = _Problem(p_start, target, sources, sub, precision, device, res)
    try:
        if method == "gauss-newton":
            mu = _gn_cg(prob, mu, steps, lam0, cg_iters, cg_tol)
        elif method == "levenberg-marquardt":
            mu = _lm(prob, mu, steps, fd_step, lam0)
        elif method == "adam":
            mu = _adam(prob, mu, steps, lr)
        elif method == "lbfgs":
            mu = _lbfgs(prob, mu, steps)
        else:
            raise ValueError(f"unknown method {method!r}")
    finally:
        prob.close()
    res.mu = mu
    return res


def _cg(matvec, b: np.ndarray, iters: int, tol: float) -> tuple[np.ndarray, int]:
    """Conjugate gradients for M x = b with M symmetric positive definite,
    from x = 0, until |r| <= tol |b| or iters products."""
    x = np.zeros_like(b)
    r = b.copy()
    pdir = r.copy()
    rr = float(r @ r)
    bb = rr
    k = 0
    while k < iters and rr > tol * tol * bb:
        q = matvec(pdir)
        alpha = rr / float(pdir @ q)
        x += alpha * pdir
        r -= alpha * q
        rr_new = float(r @ r)
        pdir = r + (rr_new / rr) * pdir
        rr = rr_new
        k += 1
    return x, k


def _gn_cg(prob: _Problem, mu: np.ndarray, steps: int, lam: float, cg_iters: int,
           cg_tol: float) -> np.ndarray:
    """Damped Gauss-Newton in theta = log mu with CG on J^T J products (two
    GPU passes each, tangent and adjoint).  lam is relative to the mean
    diagonal of J^T J, estimated from the first products."""
    r = prob.residuals(mu)
    L = float(np.sum(r * r))
    prob.res.losses.append(L)
    scale = None
    for _ in range(steps):
        g = prob.vjp(mu, r)  # J^T r = dL/dtheta / 2
        if scale is None:
            # mean diagonal of J^T J from a few random +-1 probes (E[z (M z)] = diag M)
            rng = np.random.default_rng(0)
            est = np.zeros_like(mu)
            for _k in range(4):
                z = rng.choice([-1.0, 1.0], mu.size)
                est += z * prob.vjp(mu, prob.jvp(mu, z))
            scale = max(float(np.mean(np.abs(est))) / 4.0, 1e-30)

        def matvec(v, lam_abs):
            return prob.vjp(mu, prob.jvp(mu, v)) + lam_abs * v

        while True:
            lam_abs = lam * scale
            delta, k = _cg(lambda v: matvec(v, lam_abs), -g, cg_iters, cg_tol)
            prob.res.cg_iterations.append(k)
            mu_new = _f32(mu * np.exp(delta))
            r_new = prob.residuals(mu_new)
            L_new = float(np.sum(r_new * r_new))
            if L_new < L:
                mu, r, L = mu_new, r_new, L_new
                lam = max(lam / 3.0, 1e-12)
                break
            # the residuals call moved the context to mu_new: back to mu
            prob.residuals(mu)
            lam *= 4.0
            if lam > 1e8:  # no step lowers L: converged to the arithmetic's resolution
                prob.res.losses.append(L)
                return mu
        prob.res.losses.append(L)
    return mu


def _lm(prob: _Problem, mu: np.ndarray, steps: int, h: float, lam: float) -> np.ndarray:
    r = prob.residuals(mu)
    L = float(np.sum(r * r))
    prob.res.losses.append(L)
    n = mu.size
    for _ in range(steps):
        g = prob.grad_theta(mu, r)
        # J^T J from forward differences in theta, each divided by the float32 step taken
        J = np.empty((r.size, n))
        for j in range(n):
            mj = mu.copy()
            mj[j] = _f32(np.array([mu[j] * np.exp(h)]))[0]
            J[:, j] = ((prob.residuals(mj) - r) / np.log(mj[j] / mu[j])).ravel()
        H = J.T @ J
        D = np.diag(np.diag(H))
        while True:
            delta = np.linalg.solve(H + lam * D, -0.5 * g)
            mu_new = _f32(mu * np.exp(delta))
            r_new = prob.residuals(mu_new)
            L_new = float(np.sum(r_new * r_new))
            if L_new < L:
                mu, r, L = mu_new, r_new, L_new
                lam = max(lam / 3.0, 1e-12)
                break
            lam *= 4.0
            if lam > 1e8:  # no step lowers L: converged to the arithmetic's resolution
                prob.res.losses.append(L)
                return mu
        prob.res.losses.append(L)
    return mu


def _adam(prob: _Problem, mu: np.ndarray, steps: int, lr: float) -> np.ndarray:
    theta = np.log(mu)
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)
    b1, b2, eps = 0.9, 0.999, 1e-12
    for step in range(1, steps + 1):
        mu32 = _f32(np.exp(theta))
        r = prob.residuals(mu32)
        prob.res.losses.append(float(np.sum(r * r)))
        g = prob.grad_theta(mu32, r)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        mh = m / (1 - b1 ** step)
        vh = v / (1 - b2 ** step)
        theta = theta - lr * mh / (np.sqrt(vh) + eps)
    return _f32(np.exp(theta))


def _lbfgs(prob: _Problem, mu: np.ndarray, evals: int) -> np.ndarray:
    from scipy.optimize import minimize

    def fg(theta):
        mu32 = _f32(np.exp(theta))
        r = prob.residuals(mu32)
        L = float(np.sum(r * r))
        prob.res.losses.append(L)
        return L, prob.grad_theta(mu32, r)

    out = minimize(fg, np.log(mu), jac=True, method="L-BFGS-B",
                   options=dict(maxiter=evals, maxfun=evals, ftol=0.0, gtol=0.0, maxcor=30))
    return _f32(np.exp(out.x))
