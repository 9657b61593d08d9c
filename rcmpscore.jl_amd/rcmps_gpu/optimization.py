"""Manifold geometry and the ground-state optimizer (src/optimization.jl).

metric(s, W1, W2) = 2 Re tr(W1^+ W2 (rho0 + metric_reg I))                 (:1)
precondition(s, W) = [(rho0 + precond_reg I)^{-1} (W (rho0 + metric_reg I))^+]^+   (:2-6)
retraction(s, W, alpha) = State(R + alpha W, K + i alpha/2 (W^+ R - R^+ W))       (:7-13)
finalize(s, f, g): precond_reg = max(10 eps, precond_reg_prefactor min(metric(g, g), 1))  (:18-22)

optimize_state (:123-143) runs one of the released OptimKit 0.5 algorithms the reference selects by
`method` -- "LBFGS" (LBFGS(mlbfgs)), "GD" (GradientDescent), "CG" (ConjugateGradient, Hager-Zhang
flavour) -- with the line search `ls` (default ls_default = SafeguardedLineSearch(); ls_backward
and ls_lenient as in :104-106), the metric as inner product, the preconditioner, the canonical
retraction and the identity transport (naive_transport!, :16).  `optimkit_optimize` is that
driver on any point type: it needs only fg, inner, precondition, retract and finalize.

The algorithms as OptimKit 0.5 runs them (its lbfgs.jl, gd.jl, cg.jl):
- LBFGS: with an empty memory the direction is -0.01 P g / |P g| (scale invariant); otherwise
  -H g from the two-loop recursion over the stored pairs with the preconditioner scaled by
  1 / (rho_m <y_m, P y_m>) of the newest pair m.  Pairs are stored normalised, (s/|s|, y/|s|,
  rho = |s|^2/<s,y>), only when <s,y>/<s,s> > |g|/10^4 at the new point, the oldest dropped
  beyond m.  A memory direction that is not downhill empties the memory and takes the first-step
  direction.  Every line search starts at alpha = 1 and may accept its first trial.
- GD: direction -P g; the first initial step is 1/|P g|, each next one twice the last accepted.
- CG: direction -P g + beta d_prev with the Hager-Zhang beta (eta = 0.4, theta = 1),
  beta = <y, P g>/<d, y> - theta <y, P y> <d, g> / <d, y>^2, bounded below by
  eta <d, g_prev>/<d, d> (y = g - g_prev, all inner products at the new point), beta = 0 on
  the first iteration (no periodic restart); initial steps as GD.
GD and CG searches do not accept a first trial unchecked (OptimKit's default acceptfirst = false).
finalize runs after every accepted step, not before the first.  Convergence: |g| <= gradtol;
stop after n_iter iterations.
"""
from __future__ import annotations

import math

import numpy as np

from .linesearch import BackwardLineSearch, HagerZhangLineSearch, SafeguardedLineSearch
from .observables import fg_phi4
from .state import State


def metric(s: State, w1, w2) -> float:
    reg = s.manifold_param.metric_reg
    return float(2 * np.vdot(w1, np.asarray(w2) @ (s.rho + reg * np.eye(s.D))).real)


def precondition(s: State, w):
    p = s.manifold_param
    if not p.precond_reg > 0:
        raise ValueError("preconditioner regularizer must be positive")
    rhs = (np.asarray(w) @ (s.rho + p.metric_reg * np.eye(s.D))).conj().T
    return np.linalg.solve(s.rho + p.precond_reg * np.eye(s.D), rhs).conj().T


def retraction(s: State, w, alpha: float):
    w = np.asarray(w, dtype=np.complex128)
    if w.shape != s.R.shape:
        raise ValueError("tangent has the wrong dimensions")
    k = s.K + (1j * alpha / 2) * (w.conj().T @ s.R - s.R.conj().T @ w)
    k = (k + k.conj().T) / 2
    out = State(s.R + alpha * w, k, device=s.device)
    out.copy_param(s)
    return out, w


def finalize(s: State, f, g):
    """finalize! (:18-22): precond_reg follows the gradient norm."""
    p = s.manifold_param
    p.precond_reg = max(10 * np.finfo(float).eps, p.precond_reg_prefactor * min(metric(s, g, g), 1.0))
    return s


# the line searches the reference names (src/optimization.jl:104-106)
ls_backward = BackwardLineSearch()
ls_default = SafeguardedLineSearch()
ls_lenient = HagerZhangLineSearch(c1=1e-2, c2=0.99, epsilon=1e-5)


def energy_function(s: State):
    name, c = s.model_param.name, s.model_param.coupling
    if name == "phi4":
        return lambda x: fg_phi4(x, float(c))
    raise NotImplementedError(f"model {name!r}: only phi4 is built")


def lbfgs_inverse_hessian(g, pairs, inner, precond):
    """H g of OptimKit's LBFGSInverseHessian for normalised pairs (s, y, rho): the two-loop
    recursion with the preconditioner scaled by 1 / (rho_m <y_m, P y_m>) of the newest pair."""
    alphas = [0.0] * len(pairs)
    q = g.copy()
    for i in range(len(pairs) - 1, -1, -1):
        s, y, rho = pairs[i]
        alphas[i] = rho * inner(s, q)
        q = q - alphas[i] * y
    _, y_m, rho_m = pairs[-1]
    hq = precond(q) * (1.0 / (rho_m * inner(y_m, precond(y_m))))
    for i, (s, y, rho) in enumerate(pairs):
        hq = hq + (alphas[i] - rho * inner(y, hq)) * s
    return hq


def hager_zhang_beta(g, g_prev, pg, pg_prev, d_prev, inner, eta=0.4, theta=1.0):
    """The Hager-Zhang conjugate-gradient coefficient of OptimKit's ConjugateGradient flavour."""
    y = g - g_prev
    py = pg - pg_prev
    dy = inner(d_prev, y)
    beta = inner(y, pg) / dy - theta * inner(y, py) * inner(d_prev, g) / (dy * dy)
    return max(beta, eta * inner(d_prev, g_prev) / inner(d_prev, d_prev))


def optimkit_optimize(fg, x, method, *, inner, precondition, retract, finalize, ls, n_iter, gradtol, mlbfgs=20,
                      verbosity=0):
    """OptimKit's optimize(fg, x, alg) for alg = LBFGS(mlbfgs) / GradientDescent / ConjugateGradient.
    inner(x, a, b), precondition(x, g), retract(x, p, alpha) -> (y, v), finalize(x, f, g) -> x,
    ls(fg, x, p, (f, g), initialguess, *, retract, inner, acceptfirst) -> (y, f, g, v, alpha, nfg).
    Returns (x, f, g, evaluations, history [(f, |g|)])."""
    if method not in ("LBFGS", "GD", "CG"):
        raise ValueError(f"unknown optimization method: {method}")
    f, g = fg(x)
    evaluations = 1
    gn = math.sqrt(inner(x, g, g))
    history = [(f, gn)]
    pairs = []
    alpha = 1.0
    if method != "LBFGS":
        pg = precondition(x, g)
        alpha = 1.0 / math.sqrt(inner(x, pg, pg))
    d_prev = g_prev = pg_prev = None
    it = 0
    while not (gn <= gradtol or it >= n_iter):
        if method == "LBFGS":
            eta = None
            if pairs:
                eta = -lbfgs_inverse_hessian(g, pairs, lambda a, b: inner(x, a, b), lambda v: precondition(x, v))
                if not inner(x, g, eta) < 0:
                    pairs, eta = [], None
            if eta is None:
                pg = precondition(x, g)
                eta = pg * (-0.01 / math.sqrt(inner(x, pg, pg)))
            y, f, g_new, v, step, nfg = ls(fg, x, eta, (f, g), 1.0, retract=retract, inner=inner, acceptfirst=True)
        else:
            pg = precondition(x, g)
            eta = -pg
            if method == "CG" and it > 0:
                eta = eta + hager_zhang_beta(g, g_prev, pg, pg_prev, d_prev, lambda a, b: inner(x, a, b)) * d_prev
            g_prev, pg_prev, d_prev = g, pg, eta
            y, f, g_new, v, step, nfg = ls(fg, x, eta, (f, g), alpha, retract=retract, inner=inner,
                                           acceptfirst=False)
            alpha = 2 * step
        evaluations += nfg
        it += 1
        g_old = g
        x, g = finalize(y, f, g_new), g_new
        gn = math.sqrt(inner(x, g, g))
        history.append((f, gn))
        if method == "LBFGS":
            s_ = step * v
            y_ = g - g_old
            sy, ss = inner(x, s_, y_), inner(x, s_, s_)
            if ss > 0 and sy / ss > gn / 10000:
                norm_s = math.sqrt(ss)
                pairs = (pairs + [(s_ / norm_s, y_ / norm_s, ss / sy)])[-mlbfgs:]
        if verbosity > 0:
            print(f"{method} iter {it}: E = {f:.15g}, |g| = {gn:.3e}")
    return x, f, g, evaluations, history


def optimize_state(s: State, *, method: str = "LBFGS", n_iter: int = 1000, gradtol: float = 1e-7,
                   mlbfgs: int = 20, ls=None, verbosity: int = 0, diagnostics: bool = False):
    """Optimize with the released OptimKit algorithm `method` ("LBFGS", "GD" or "CG") over the
    canonical retraction with the Riemannian metric and preconditioner (src/optimization.jl:123-143)."""
    if n_iter < 0 or not gradtol > 0:
        raise ValueError("invalid optimization controls")
    if method not in ("LBFGS", "GD", "CG"):
        raise ValueError(f"unknown optimization method: {method}")
    x, f, g, evaluations, history = optimkit_optimize(
        energy_function(s), s, method, inner=metric, precondition=precondition, retract=retraction,
        finalize=finalize, ls=ls or ls_default, n_iter=n_iter, gradtol=gradtol, mlbfgs=mlbfgs, verbosity=verbosity)
    x.e0, x.grad = f, g
    gn = math.sqrt(metric(x, g, g))
    if diagnostics:
        return dict(state=x, energy=f, gradient_norm=gn, converged=gn <= gradtol, evaluations=evaluations,
                    history=history)
    return x
