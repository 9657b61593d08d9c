"""Manifold geometry and the ground-state optimizer (src/optimization.jl).

metric(s, W1, W2) = 2 Re tr(W1^+ W2 (rho0 + metric_reg I))                 (:1)
precondition(s, W) = [(rho0 + precond_reg I)^{-1} (W (rho0 + metric_reg I))^+]^+   (:2-6)
retraction(s, W, alpha) = State(R + alpha W, K + i alpha/2 (W^+ R - R^+ W))       (:7-13)
optimize_state: L-BFGS (OptimKit LBFGS(m)) with the line search ls_default =
SafeguardedLineSearch() (:54-106): Hager-Zhang with a bounded step and Armijo backtracking
(BackwardLineSearch, :24-53) as recovery; failed trials (SolverFailure) are rejected.
"""
from __future__ import annotations

import math

import numpy as np

from .linesearch import SafeguardedLineSearch
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


def energy_function(s: State):
    name, c = s.model_param.name, s.model_param.coupling
    if name == "phi4":
        return lambda x: fg_phi4(x, float(c))
    raise NotImplementedError(f"model {name!r}: only phi4 is built")


def optimize_state(s: State, *, method: str = "LBFGS", n_iter: int = 1000, gradtol: float = 1e-7,
                   mlbfgs: int = 20, ls=None, verbosity: int = 0, diagnostics: bool = False):
    """L-BFGS over the canonical retraction with the Riemannian metric and preconditioner.
    Transport is the identity (naive_transport!, :16)."""
    if n_iter < 0 or not gradtol > 0:
        raise ValueError("invalid optimization controls")
    if method not in ("LBFGS", "GD"):
        raise ValueError(f"unknown optimization method: {method}")
    fg = energy_function(s)
    ls = ls or SafeguardedLineSearch()
    x = s
    f, g = fg(x)
    finalize(x, f, g)
    history = [(f, math.sqrt(metric(x, g, g)))]
    evaluations = 1
    S, Y, RHO = [], [], []
    for it in range(n_iter):
        gn = math.sqrt(metric(x, g, g))
        if gn <= gradtol:
            break
        # two-loop recursion in the metric of the current point
        q = g.copy()
        alphas = []
        if method == "LBFGS":
            for sk, yk, rk in reversed(list(zip(S, Y, RHO))):
                a = rk * metric(x, sk, q)
                alphas.append(a)
                q = q - a * yk
        r = precondition(x, q)
        if method == "LBFGS" and S:
            for (sk, yk, rk), a in zip(zip(S, Y, RHO), reversed(alphas)):
                b = rk * metric(x, yk, r)
                r = r + (a - b) * sk
        p = -r
        if metric(x, g, p) >= 0:
            p = -precondition(x, g)
            S, Y, RHO = [], [], []
        y, fy, gy, v, alpha, nfg = ls(fg, x, p, (f, g), retract=retraction, inner=metric)
        evaluations += nfg
        sk = alpha * v
        yk = gy - g
        sy = metric(y, sk, yk)
        if method == "LBFGS" and sy > 0:
            S.append(sk)
            Y.append(yk)
            RHO.append(1.0 / sy)
            if len(S) > mlbfgs:
                S.pop(0), Y.pop(0), RHO.pop(0)
        x, f, g = y, fy, gy
        finalize(x, f, g)
        history.append((f, math.sqrt(metric(x, g, g))))
        if verbosity > 0:
            print(f"iter {it + 1}: E = {f:.15g}, |g| = {history[-1][1]:.3e}")
    x.e0, x.grad = f, g
    gn = math.sqrt(metric(x, g, g))
    if diagnostics:
        return dict(state=x, energy=f, gradient_norm=gn, converged=gn <= gradtol, evaluations=evaluations,
                    history=history)
    return x
