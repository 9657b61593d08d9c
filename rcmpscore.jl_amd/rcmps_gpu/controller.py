"""Resumable L-BFGS controller over (R, K) for long optimizations, with the resolution guards and
the two line searches of the reference GPU package's runtime controller
(gpu/src/rcmps_gpu/optimizer/runtime/optimizer.py:23-259, robust_optimizer.py:30-505).

`optimize` is a generator of events on an already evaluated accepted state:

- {"type": "trial", ...} for every line-search trial (step, energy, acceptance, failures);
- {"type": "accepted", "state": snapshot, ...} after every accepted step;
- {"type": "stopped", "reason": ..., "converged": bool, "state": snapshot} once, at the end.

Every snapshot holds what an exact continuation needs: R, K, rho, gradient, energy, the L-BFGS
pairs normalised to unit step norm (s/|s|, y/|s|, |s|^2/<s,y>), precond_reg, the next initial
step, the accepted-iteration and evaluation counters and the elapsed seconds.  Passing it back as
`resume_state` continues the run with no objective call for the starting point:
`max_iterations` caps the accepted iterations across resumes, `max_seconds` only this call.
`save_resume_state` / `load_resume_state` keep a snapshot in one .npz file (no pickles).

The step direction is the preconditioned two-loop L-BFGS direction in the metric
2 Re tr(x^+ y (rho + metric_reg I)), scaled by the latest secant (OptimKit 0.5's scaling of the
preconditioner), with the cautious update <s,y>/|s|^2 > |g|/10^4; the retraction is the canonical
one, K + i alpha/2 (W^+ R - R^+ W) (src/optimization.jl:7-13).  Without `robustness` the line
search is the bounded Armijo backtracking of the reference controller.  With a
RobustnessControls it adds the guards of the robust controller: trials whose predicted decrease
or step are below the energy / step resolution are not evaluated, secants that the energy and
gradient do not resolve are not stored, and a failed search is retried at most max_recoveries
times (first with the history cleared, then along the plain steepest direction) before the run
reports numerical stagnation; line_search="wolfe" runs the Hager-Zhang search of
SafeguardedLineSearch first, with Armijo backtracking as its fallback.

`stationary(q, r, tolerance) -> (rho, stats)` solves the stationary density of a trial (on the
GPU by default, through State); `objective(r, k, rho) -> (energy, gradient, stats)` evaluates it
(fg_phi4 on the GPU for `gpu_objective`).  Both are arguments, so the controller itself is host
logic and runs with any evaluator.
"""
from __future__ import annotations

import json
import math
import os
import tempfile
from dataclasses import dataclass, fields
from pathlib import Path
from time import perf_counter

import numpy as np

from ._lib import SolverFailure
from .linesearch import HagerZhangLineSearch, Trial

EPS = float(np.finfo(float).eps)


@dataclass(frozen=True)
class OptimizationControls:
    """Controls of optimize (runtime/optimizer.py:23-40 names and defaults)."""
    max_iterations: int = 3000
    gradient_tol: float = 3e-5
    max_seconds: float | None = None
    history: int = 100
    metric_reg: float = 1e-3
    stationary_tolerance: float = 1e-13
    initial_precond_reg: float = 1e-4
    precond_prefactor: float = 1e-2
    initial_step: float = 1.0
    initial_direction_norm: float = 0.01
    warm_start_line_search: bool = False
    max_relative_step: float = 0.5
    line_search_trials: int = 30
    armijo_c1: float = 1e-4
    shrink: float = 0.5
    energy_atol: float = 0.0
    refresh_curvatures: bool = False


_ALIASES = {"iterations": "max_iterations", "gradtol": "gradient_tol", "gradient_tolerance": "gradient_tol",
            "lbfgs_history": "history", "max_trials": "line_search_trials", "armijo_energy_atol": "energy_atol"}


def _finite_positive(x) -> bool:
    return x is not None and math.isfinite(x) and x > 0


def check_controls(controls=None) -> OptimizationControls:
    """An OptimizationControls from itself, a dict or an object with attributes (older names are
    accepted as aliases, unknown ones ignored); ValueError for invalid values
    (runtime/optimizer.py:43-67)."""
    if controls is None:
        c = OptimizationControls()
    elif isinstance(controls, OptimizationControls):
        c = controls
    else:
        values = dict(controls) if isinstance(controls, dict) else dict(vars(controls))
        for old, new in _ALIASES.items():
            if old in values and new not in values:
                values[new] = values[old]
        known = {f.name for f in fields(OptimizationControls)}
        c = OptimizationControls(**{k: v for k, v in values.items() if k in known})
    ok = (c.max_iterations >= 0 and c.history >= 1 and c.line_search_trials >= 1
          and _finite_positive(c.gradient_tol) and math.isfinite(c.metric_reg) and c.metric_reg >= 0
          and all(_finite_positive(v) for v in (c.stationary_tolerance, c.initial_precond_reg, c.precond_prefactor,
                                                 c.initial_step, c.initial_direction_norm, c.max_relative_step))
          and 0 < c.armijo_c1 < 1 and 0 < c.shrink < 1 and math.isfinite(c.energy_atol) and c.energy_atol >= 0
          and (c.max_seconds is None or _finite_positive(c.max_seconds)))
    if not ok:
        raise ValueError("invalid optimization controls")
    return c


@dataclass(frozen=True)
class RobustnessControls:
    """Resolution guards and line search of the robust controller (robust_optimizer.py:30-52)."""
    energy_resolution_factor: float = 64.0
    step_resolution_factor: float = 64.0
    predicted_decrease_factor: float = 10.0
    secant_gradient_floor: float = 1e-10
    max_recoveries: int = 2
    line_search: str = "armijo"
    wolfe_energy_atol: float | None = None

    def __post_init__(self):
        if (not all(_finite_positive(v) for v in (self.energy_resolution_factor, self.step_resolution_factor,
                                                  self.predicted_decrease_factor))
                or not math.isfinite(self.secant_gradient_floor) or self.secant_gradient_floor < 0
                or self.max_recoveries not in (1, 2)):
            raise ValueError("invalid robust optimization controls")
        wolfe = self.line_search == "wolfe"
        if (self.line_search not in ("armijo", "wolfe") or wolfe != (self.wolfe_energy_atol is not None)
                or (wolfe and not (math.isfinite(self.wolfe_energy_atol) and self.wolfe_energy_atol >= 0))):
            raise ValueError("invalid line search controls")


# ---------------------------------------------------------------------------------------------
# geometry on plain arrays (runtime/numerics.py:129-142, src/optimization.jl:1-13)
# ---------------------------------------------------------------------------------------------
def metric(rho, x, y, metric_reg: float) -> float:
    """2 Re tr(x^+ y (rho + metric_reg I))."""
    return float(2 * np.vdot(x, y @ (rho + metric_reg * np.eye(len(rho)))).real)


def precondition(rho, g, metric_reg: float, precond_reg: float):
    eye = np.eye(len(rho))
    return np.linalg.solve(rho + precond_reg * eye, (g @ (rho + metric_reg * eye)).conj().T).conj().T


def retract(r, k, direction, alpha: float):
    """(R + alpha W, K + i alpha/2 (W^+ R - R^+ W)), K re-Hermitised."""
    rr = r + alpha * direction
    kk = k + 0.5j * alpha * (direction.conj().T @ r - r.conj().T @ direction)
    return rr, (kk + kk.conj().T) / 2


def _norm(rho, x, metric_reg: float) -> float:
    sq = metric(rho, x, x, metric_reg)
    if not (math.isfinite(sq) and sq >= 0):
        raise SolverFailure(f"invalid tangent metric norm squared: {sq}")
    return math.sqrt(sq)


def _steepest(rho, g, c: OptimizationControls, precond_reg: float):
    """The preconditioned gradient scaled to initial_direction_norm, negated."""
    pg = precondition(rho, g, c.metric_reg, precond_reg)
    n = _norm(rho, pg, c.metric_reg)
    if n == 0:
        raise SolverFailure("preconditioned direction vanished before convergence")
    return -(c.initial_direction_norm / n) * pg


def _lbfgs(rho, g, pairs, c: OptimizationControls, precond_reg: float):
    """Two-loop recursion with the preconditioner rescaled to the latest secant; (direction, gamma),
    direction None when the scaling is not positive."""
    q = g.copy()
    coef = []
    for s, y, inv in reversed(pairs):
        a = inv * metric(rho, s, q, c.metric_reg)
        coef.append(a)
        q = q - a * y
    s_last, y_last, _ = pairs[-1]
    num = metric(rho, s_last, y_last, c.metric_reg)
    den = metric(rho, y_last, precondition(rho, y_last, c.metric_reg, precond_reg), c.metric_reg)
    gamma = num / den if den > 0 else math.nan
    if not (math.isfinite(gamma) and gamma > 0):
        return None, gamma
    d = gamma * precondition(rho, q, c.metric_reg, precond_reg)
    for (s, y, inv), a in zip(pairs, reversed(coef)):
        d = d + (a - inv * metric(rho, y, d, c.metric_reg)) * s
    return -d, float(gamma)


def _wolfe_check(t: Trial, origin: Trial, eps: float, hz: HagerZhangLineSearch):
    """The returned trial's condition: "exact", "approximate" or None (robust_optimizer.py:100-112)."""
    if not (math.isfinite(t.f) and math.isfinite(t.slope) and t.alpha > 0):
        return None
    if t.f - origin.f <= hz.c1 * t.alpha * origin.slope and t.slope >= hz.c2 * origin.slope:
        return "exact"
    if t.f <= origin.f + eps and hz.c2 * origin.slope <= t.slope <= (2 * hz.c1 - 1) * origin.slope:
        return "approximate"
    return None


class _SearchEnd(Exception):
    pass


class _OutOfTime(Exception):
    pass


def optimize(r, k, rho, energy, gradient, objective, controls=None, resume_state=None, *, robustness=None,
             stationary=None):
    """Yield trial / accepted / stopped events from the accepted state (r, k, rho, energy, gradient)
    or from `resume_state` (a snapshot of an earlier run, which then overrides them)."""
    c = check_controls(controls)
    if stationary is None:
        stationary = gpu_stationary()
    rs = {} if resume_state is None else resume_state
    r = np.array(rs.get("R", r), dtype=np.complex128)
    k = np.array(rs.get("K", k), dtype=np.complex128)
    rho = np.array(rs.get("rho", rho), dtype=np.complex128)
    gradient = np.array(rs.get("gradient", gradient), dtype=np.complex128)
    energy = float(rs.get("energy", energy))
    if (r.ndim != 2 or r.shape[0] != r.shape[1] or any(a.shape != r.shape for a in (k, rho, gradient))
            or not math.isfinite(energy) or not all(np.isfinite(a).all() for a in (r, k, rho, gradient))):
        raise ValueError("invalid accepted optimization state")
    pairs = [(np.array(s, dtype=np.complex128), np.array(y, dtype=np.complex128), float(inv))
             for s, y, inv in rs.get("pairs", [])][-c.history:]
    if any(s.shape != r.shape or y.shape != r.shape or not (math.isfinite(inv) and inv > 0)
           or not np.isfinite(s).all() or not np.isfinite(y).all() for s, y, inv in pairs):
        raise ValueError("invalid resumed LBFGS history")
    iteration = int(rs.get("iteration", 0))
    evaluations = int(rs.get("evaluations", 0))
    precond_reg = float(rs.get("precond_reg", c.initial_precond_reg))
    next_step = float(rs.get("next_step", c.initial_step))
    elapsed_before = float(rs.get("elapsed_seconds", 0.0))
    if (iteration < 0 or evaluations < 0 or not _finite_positive(precond_reg) or not _finite_positive(next_step)
            or not (math.isfinite(elapsed_before) and elapsed_before >= 0)):
        raise ValueError("invalid resumed optimizer controls or counters")
    guards = robustness
    wolfe = guards is not None and guards.line_search == "wolfe"
    hz = HagerZhangLineSearch(epsilon=guards.wolfe_energy_atol if wolfe else 1e-6)
    last_stats = rs.get("stats")
    began = perf_counter()
    recovery = 0
    unresolved_run = 0

    def snapshot():
        return dict(R=r, K=k, rho=rho, gradient=gradient, energy=energy, pairs=list(pairs),
                    precond_reg=precond_reg, next_step=next_step, iteration=iteration, evaluations=evaluations,
                    elapsed_seconds=elapsed_before + perf_counter() - began,
                    gradient_norm=_norm(rho, gradient, c.metric_reg), stats=last_stats)

    def stopped(reason, converged=False, **extra):
        return dict(type="stopped", reason=reason, converged=converged, state=snapshot(), **extra)

    def out_of_time():
        return c.max_seconds is not None and perf_counter() - began >= c.max_seconds

    def evaluate(direction, alpha):
        """(rr, kk, rho, energy, gradient, stats) at step alpha; SolverFailure / RuntimeError pass."""
        nonlocal evaluations
        rr, kk = retract(r, k, direction, alpha)
        q = -1j * kk - rr.conj().T @ rr / 2
        rnew, sstats = stationary(q, rr, c.stationary_tolerance)
        evaluations += 1
        e, g, st = objective(rr, kk, rnew)
        e = float(e)
        g = np.asarray(g, dtype=np.complex128)
        if not math.isfinite(e) or g.shape != r.shape or not np.isfinite(g).all():
            raise SolverFailure("trial returned nonfinite or invalid objective/gradient")
        return rr, kk, np.asarray(rnew, dtype=np.complex128), e, g, st, sstats

    while True:
        gn = _norm(rho, gradient, c.metric_reg)
        if gn <= c.gradient_tol:
            yield stopped("gradient tolerance reached", converged=True)
            return
        if iteration >= c.max_iterations:
            yield stopped("requested iteration limit reached")
            return
        if out_of_time():
            yield stopped("optimization time budget reached")
            return
        t_iter = perf_counter()
        e_res = max(c.energy_atol, guards.energy_resolution_factor * EPS * max(1.0, abs(energy))) if guards else 0.0
        s_res = guards.step_resolution_factor * EPS * max(1.0, np.linalg.norm(r)) if guards else 0.0
        if c.refresh_curvatures:
            kept = []
            for s, y, _ in pairs:
                sy, ss = metric(rho, s, y, c.metric_reg), metric(rho, s, s, c.metric_reg)
                if ss > 0 and sy / ss > gn / 10000:
                    kept.append((s, y, 1 / sy))
            pairs = kept
        gamma = None
        if guards is not None and recovery == 2:
            direction = -(c.initial_direction_norm / gn) * gradient
        elif pairs:
            direction, gamma = _lbfgs(rho, gradient, pairs, c, precond_reg)
        else:
            direction = _steepest(rho, gradient, c, precond_reg)
        slope = metric(rho, gradient, direction, c.metric_reg) if direction is not None else math.nan
        if not (math.isfinite(slope) and slope < 0):
            pairs = []
            direction = _steepest(rho, gradient, c, precond_reg)
            slope = metric(rho, gradient, direction, c.metric_reg)
        if not (math.isfinite(slope) and slope < 0):
            raise SolverFailure("preconditioned direction is not a descent direction")
        alpha_max = c.max_relative_step * max(1.0, np.linalg.norm(r)) / max(np.linalg.norm(direction), EPS)
        alpha = min(next_step, alpha_max)
        accepted, failures, unresolved_trial, trial = False, [], False, 0
        found = None
        searched, tag = {}, {}
        expected = -alpha * slope
        if wolfe:
            origin = Trial(0.0, energy, float(slope))
            log = []

            def take(step):
                if out_of_time():
                    raise _OutOfTime
                if step > alpha_max:
                    raise _SearchEnd(f"trial step {step:.3e} exceeds the step bound {alpha_max:.3e}")
                if step * np.linalg.norm(direction) <= s_res:
                    raise _SearchEnd(f"trial step {step:.3e} is below the step resolution")
                try:
                    out = evaluate(direction, step)
                except (SolverFailure, RuntimeError) as err:
                    failures.append(str(err))
                    log.append(dict(type="trial", iteration=iteration + 1, trial=len(log) + 1, line_search="wolfe",
                                    alpha=float(step), accepted=False, error=str(err), evaluations=evaluations))
                    raise _SearchEnd(str(err)) from None
                t = Trial(float(step), out[3], metric(out[2], out[4], direction, c.metric_reg), out)
                exact, approx = hz.conditions(t, origin)
                log.append(dict(type="trial", iteration=iteration + 1, trial=len(log) + 1, line_search="wolfe",
                                alpha=t.alpha, energy=t.f, slope=t.slope, initial_slope=origin.slope,
                                exact_wolfe=exact, approximate_wolfe=approx, accepted=False, stats=out[5],
                                evaluations=evaluations))
                return t

            condition = reason = None
            try:
                t, _ = hz.search(take, origin, alpha, acceptfirst=True)
                condition = _wolfe_check(t, origin, guards.wolfe_energy_atol, hz)
                if condition is None:
                    reason = f"the search ended at step {t.alpha:.3e} without the Wolfe conditions"
            except _SearchEnd as err:
                reason = str(err)
            except SolverFailure as err:
                failures.append(str(err))
                reason = str(err)
            except _OutOfTime:
                yield from log
                yield stopped("optimization time budget reached during line search", numerical_failures=failures)
                return
            yield from log
            yield dict(type="trial", iteration=iteration + 1, trial=len(log), line_search="wolfe", stage="result",
                       alpha=t.alpha if condition else None, accepted=condition is not None,
                       wolfe_condition=condition, armijo_fallback=condition is None, reason=reason,
                       evaluations=evaluations)
            searched = dict(line_search="wolfe", wolfe_condition=condition, wolfe_trials=len(log),
                            wolfe_failure=reason, armijo_fallback=condition is None)
            tag = dict(line_search="armijo")
            if condition is not None:
                found, alpha, accepted, trial = t.data, t.alpha, True, len(log)
                expected = -alpha * slope
        if not accepted:
            for trial in range(1, c.line_search_trials + 1):
                if out_of_time():
                    yield stopped("optimization time budget reached during line search", numerical_failures=failures)
                    return
                expected = -alpha * slope
                nominal = float(alpha * np.linalg.norm(direction))
                if guards is not None and (expected <= guards.predicted_decrease_factor * e_res or nominal <= s_res):
                    unresolved_trial = True
                    yield dict(type="trial", iteration=iteration + 1, trial=trial, alpha=float(alpha), accepted=False,
                               robust_skipped_unresolved=True, expected_decrease=float(expected),
                               energy_resolution=float(e_res), nominal_step_norm=nominal,
                               step_resolution=float(s_res), recovery_mode=recovery, evaluations=evaluations, **tag)
                    break
                try:
                    out = evaluate(direction, alpha)
                    bound = energy + c.armijo_c1 * alpha * slope + c.energy_atol
                    accepted = bool(out[3] <= bound)
                    yield dict(type="trial", iteration=iteration + 1, trial=trial, alpha=float(alpha), energy=out[3],
                               armijo_bound=float(bound), accepted=accepted, stats=out[5], evaluations=evaluations,
                               **tag)
                    if accepted:
                        found = out
                        break
                except (SolverFailure, RuntimeError) as err:
                    failures.append(str(err))
                    yield dict(type="trial", iteration=iteration + 1, trial=trial, alpha=float(alpha), accepted=False,
                               error=str(err), evaluations=evaluations, **tag)
                alpha *= c.shrink
            if wolfe:
                trial += searched["wolfe_trials"]
        if not accepted:
            if guards is None:
                yield stopped("Armijo line search exhausted; last accepted state retained", numerical_failures=failures)
                return
            if recovery < guards.max_recoveries:
                pairs, next_step = [], c.initial_step
                recovery += 1
                yield dict(type="trial", iteration=iteration + 1, trial=trial, alpha=float(next_step), accepted=False,
                           robust_restart=True, reason="unresolved trial" if unresolved_trial else "line search exhausted",
                           recovery_mode=recovery, history_size=0, evaluations=evaluations)
                continue
            if unresolved_trial:
                yield stopped("Numerical stagnation: unresolved trial after bounded clean recovery",
                              numerical_failures=failures, recovery_mode=recovery)
                return
            yield stopped(("Wolfe search and Armijo backtracking" if wolfe else "Armijo line search")
                          + " exhausted; last accepted state retained", numerical_failures=failures,
                          recovery_mode=recovery)
            return
        rr, kk, rnew, enew, gnew, st_new, _ = found
        step, change = alpha * direction, gnew - gradient
        new_gn = _norm(rnew, gnew, c.metric_reg)
        sy, ss = metric(rnew, step, change, c.metric_reg), metric(rnew, step, step, c.metric_reg)
        extra = {}
        resolved = True
        if guards is not None:
            dg = _norm(rnew, change, c.metric_reg)
            dr = float(np.linalg.norm(rr - r))
            decrease = energy - enew
            if wolfe:
                # the gradient resolves the secant where the energy cannot (OptimKit's cautious rule decides)
                resolved = dr > s_res and dg > guards.secant_gradient_floor
            else:
                resolved = (decrease > e_res and expected > guards.predicted_decrease_factor * e_res
                            and dr > s_res and dg > guards.secant_gradient_floor)
            extra = dict(robust_secant_resolved=bool(resolved), actual_step_norm=dr, step_resolution=float(s_res),
                         gradient_change_norm=dg, actual_energy_decrease=float(decrease),
                         expected_decrease=float(expected), energy_resolution=float(e_res))
        cautious = resolved and math.isfinite(sy) and math.isfinite(ss) and ss > 0 and sy / ss > new_gn / 10000
        if cautious:
            ns = math.sqrt(ss)
            pairs = (pairs + [(step / ns, change / ns, ss / sy)])[-c.history:]
        r, k, rho, energy, gradient = rr, kk, rnew, enew, gnew
        last_stats = st_new
        iteration += 1
        next_step = min(1.0, 2 * alpha) if c.warm_start_line_search else c.initial_step
        precond_reg = max(10 * EPS, c.precond_prefactor * min(new_gn * new_gn, 1.0))
        stop_unresolved = False
        if guards is not None:
            unresolved = (extra["actual_step_norm"] <= s_res
                          or (extra["actual_energy_decrease"] <= e_res
                              and extra["gradient_change_norm"] <= guards.secant_gradient_floor))
            if unresolved:
                pairs, next_step = [], c.initial_step
                unresolved_run += 1
                stop_unresolved = unresolved_run > guards.max_recoveries
                recovery = min(guards.max_recoveries, recovery + 1)
            else:
                unresolved_run, recovery = 0, 0
            extra.update(robust_unresolved_accepted=bool(unresolved), recovery_mode=recovery)
        yield dict(type="accepted", state=snapshot(), iteration=iteration, energy=energy, gradient_norm=new_gn,
                   alpha=float(alpha), trials=trial, seconds=perf_counter() - t_iter, stats=st_new, gamma=gamma,
                   cautious_update=bool(cautious), history_size=len(pairs), evaluations=evaluations, **extra,
                   **searched)
        if stop_unresolved and new_gn > c.gradient_tol:
            yield stopped("Numerical stagnation: repeated unresolved accepted steps after bounded clean recovery",
                          unresolved_acceptances=unresolved_run)
            return


# ---------------------------------------------------------------------------------------------
# evaluators on the GPU
# ---------------------------------------------------------------------------------------------
class _StateCache:
    """The last State solved by gpu_stationary, so that gpu_objective reuses it for the same trial."""
    state = None


def gpu_stationary(device: int = 0):
    """stationary(q, r, tolerance) -> (rho, stats): the stationary density of the trial (R, Q) solved
    on the GPU (State, src/state.jl:40-53); K = i (Q + R^+R/2)."""
    from .state import State

    def solve(q, r, tolerance):
        k = 1j * (q + r.conj().T @ r / 2)
        k = (k + k.conj().T) / 2
        s = State(r, k, steady_tol=tolerance, device=device)
        _StateCache.state = s
        return s.rho, None
    return solve


def gpu_objective(coupling: float, *, integrator=None, device: int = 0, precision: str = "fp64"):
    """objective(r, k, rho) -> (energy, gradient, stats): fg_phi4 at this coupling on the GPU; the
    State is reused from gpu_stationary when it is the same trial, else built with rho checked."""
    from .observables import fg_phi4
    from .state import State

    def evaluate(r, k, rho):
        s = _StateCache.state
        if s is None or not (np.array_equal(s.R, r) and np.array_equal(s.K, k)):
            s = State(r, k, rho=rho, device=device)
        if integrator is not None:
            s.integrator_param = integrator
        e, g, st = fg_phi4(s, float(coupling), return_stats=True, precision=precision)
        return e, g, st
    return evaluate


# ---------------------------------------------------------------------------------------------
# snapshots on disk
# ---------------------------------------------------------------------------------------------
_SCALARS = ("energy", "precond_reg", "next_step", "iteration", "evaluations", "elapsed_seconds")


def save_resume_state(path, state) -> Path:
    """Write a snapshot (arrays, normalised L-BFGS pairs, counters) to one .npz file through a
    temporary file and an atomic rename; no pickles."""
    path = Path(path)
    pairs = state.get("pairs", [])
    d = np.asarray(state["R"]).shape[0]
    meta = {key: state[key] for key in _SCALARS if key in state}
    meta["schema"] = "rcmps_gpu resume state 1"
    arrays = dict(R=np.asarray(state["R"], dtype=np.complex128), K=np.asarray(state["K"], dtype=np.complex128),
                  rho=np.asarray(state["rho"], dtype=np.complex128),
                  gradient=np.asarray(state["gradient"], dtype=np.complex128),
                  pair_s=np.array([p[0] for p in pairs], dtype=np.complex128).reshape(len(pairs), d, d),
                  pair_y=np.array([p[1] for p in pairs], dtype=np.complex128).reshape(len(pairs), d, d),
                  pair_inv=np.array([p[2] for p in pairs], dtype=np.float64),
                  meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    path.parent.mkdir(parents=True, exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=path.parent, suffix=".npz.tmp")
    os.close(fd)
    try:
        with open(tmp, "wb") as f:
            np.savez(f, **arrays)
        os.replace(tmp, path)
    finally:
        if os.path.exists(tmp):
            os.remove(tmp)
    return path


def load_resume_state(path) -> dict:
    """A snapshot written by save_resume_state, ready for optimize(resume_state=...)."""
    with np.load(Path(path), allow_pickle=False) as z:
        meta = json.loads(bytes(z["meta"]).decode())
        if meta.get("schema") != "rcmps_gpu resume state 1":
            raise ValueError("not an rcmps_gpu resume state")
        state = {key: z[key].copy() for key in ("R", "K", "rho", "gradient")}
        state["pairs"] = [(z["pair_s"][i].copy(), z["pair_y"][i].copy(), float(z["pair_inv"][i]))
                          for i in range(len(z["pair_inv"]))]
    for key in _SCALARS:
        if key in meta:
            state[key] = meta[key]
    return state
