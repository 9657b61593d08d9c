"""Resumable L-BFGS optimization over (R, K) for long runs, with resolution guards and a choice
of line search.

This serves the callers of the reference GPU package's runtime controllers
(gpu/src/rcmps_gpu/optimizer/runtime/optimizer.py, robust_optimizer.py): the controls keep their
names and defaults, the events their kinds ("trial", "accepted", "stopped") and the snapshots their
fields, so a caller finds the same knobs, stops and resume semantics.  The organisation is this
module's own:

* `_Geometry` holds the metric 2 Re tr(x^+ y (rho + metric_reg I)) of one density, its norm and the
  preconditioner (src/optimization.jl:1-13);
* `_Memory` is the L-BFGS secant memory.  Pairs are kept scaled to unit step length
  (s/|s|, y/|s|, |s|^2/<s,y>), so a snapshot is independent of the step sizes that made it; the
  two-loop recursion scales the preconditioner by the newest pair's <s,y>/<y,P y> (OptimKit 0.5's
  scaling), and a pair enters only when <s,y>/|s|^2 exceeds |g|/10^4 (the cautious rule);
* `_Guard` is the guarded mode (a RobustnessControls): energy and step floors under which a trial is
  noise, and a recovery level that rises after a failed search or an accepted step the floors do not
  resolve (level 1: memory dropped; level 2: the plain gradient direction), and falls back to 0 after
  a resolved step.  A search that fails at the top level ends the run;
* `_Searcher` runs one line search along a direction and returns the accepted `_Point` or why none
  was found: bounded Armijo backtracking, or (line_search="wolfe") the Hager-Zhang search of
  SafeguardedLineSearch whose result must meet the exact or approximate Wolfe condition, with
  backtracking from the same first step as its fallback;
* `optimize` is the generator that runs them from an evaluated point or a snapshot.

Events: {"type": "trial", ...} per line-search trial (step, energy, acceptance, failure),
{"type": "accepted", "state": snapshot, ...} per accepted step, and one {"type": "stopped",
"reason", "converged", "state"} at the end.  A snapshot holds R, K, rho, gradient, energy, the
memory pairs, precond_reg, next_step, iteration, evaluations, elapsed_seconds, gradient_norm and
the last stats; passed back as `resume_state` it continues the run exactly, with no evaluation of
its starting point (`max_iterations` counts across resumes, `max_seconds` per call).
`save_resume_state` / `load_resume_state` keep one in an .npz file without pickles.

`stationary(q, r, tolerance) -> (rho, stats)` solves a trial's density and
`objective(r, k, rho) -> (energy, gradient, stats)` evaluates it; `gpu_stationary` and
`gpu_objective` do both on the GPU (State and fg_phi4).
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


# ---------------------------------------------------------------------------------------------
# controls
# ---------------------------------------------------------------------------------------------
@dataclass(frozen=True)
class OptimizationControls:
    """Controls of optimize; names and defaults as the reference runtime's controller."""
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

    def problems(self):
        """Names of the controls whose values are invalid."""
        positive = ("gradient_tol", "stationary_tolerance", "initial_precond_reg", "precond_prefactor",
                    "initial_step", "initial_direction_norm", "max_relative_step")
        bad = [name for name in positive if not _positive(getattr(self, name))]
        if self.max_iterations < 0:
            bad.append("max_iterations")
        if self.history < 1:
            bad.append("history")
        if self.line_search_trials < 1:
            bad.append("line_search_trials")
        if not (math.isfinite(self.metric_reg) and self.metric_reg >= 0):
            bad.append("metric_reg")
        if not 0 < self.armijo_c1 < 1:
            bad.append("armijo_c1")
        if not 0 < self.shrink < 1:
            bad.append("shrink")
        if not (math.isfinite(self.energy_atol) and self.energy_atol >= 0):
            bad.append("energy_atol")
        if self.max_seconds is not None and not _positive(self.max_seconds):
            bad.append("max_seconds")
        return bad


# earlier names of some controls, still accepted
_RENAMED = {"iterations": "max_iterations", "gradtol": "gradient_tol", "gradient_tolerance": "gradient_tol",
            "lbfgs_history": "history", "max_trials": "line_search_trials", "armijo_energy_atol": "energy_atol"}


def _positive(x) -> bool:
    return x is not None and math.isfinite(x) and x > 0


def check_controls(controls=None) -> OptimizationControls:
    """An OptimizationControls from None, itself, a dict or an object with attributes (earlier names
    accepted, unknown names ignored); ValueError naming the invalid controls."""
    if controls is None:
        c = OptimizationControls()
    elif isinstance(controls, OptimizationControls):
        c = controls
    else:
        given = dict(controls) if isinstance(controls, dict) else dict(vars(controls))
        for old, new in _RENAMED.items():
            if old in given and new not in given:
                given[new] = given[old]
        names = {f.name for f in fields(OptimizationControls)}
        c = OptimizationControls(**{key: value for key, value in given.items() if key in names})
    bad = c.problems()
    if bad:
        raise ValueError("invalid optimization controls: " + ", ".join(bad))
    return c


@dataclass(frozen=True)
class RobustnessControls:
    """The guarded mode: floors of energy and step resolution (factors of machine epsilon), the
    predicted-decrease margin, the gradient-change floor of a stored secant, the number of recovery
    levels (1 or 2) and the line search ("armijo", or "wolfe" with its approximate-Wolfe energy
    tolerance)."""
    energy_resolution_factor: float = 64.0
    step_resolution_factor: float = 64.0
    predicted_decrease_factor: float = 10.0
    secant_gradient_floor: float = 1e-10
    max_recoveries: int = 2
    line_search: str = "armijo"
    wolfe_energy_atol: float | None = None

    def __post_init__(self):
        factors = (self.energy_resolution_factor, self.step_resolution_factor, self.predicted_decrease_factor)
        if (not all(_positive(f) for f in factors) or self.max_recoveries not in (1, 2)
                or not (math.isfinite(self.secant_gradient_floor) and self.secant_gradient_floor >= 0)):
            raise ValueError("invalid robust optimization controls")
        if self.line_search == "armijo":
            ok = self.wolfe_energy_atol is None
        elif self.line_search == "wolfe":
            tol = self.wolfe_energy_atol
            ok = tol is not None and math.isfinite(tol) and tol >= 0
        else:
            ok = False
        if not ok:
            raise ValueError("invalid line search controls")


# ---------------------------------------------------------------------------------------------
# geometry of one density (src/optimization.jl:1-13)
# ---------------------------------------------------------------------------------------------
def metric(rho, x, y, metric_reg: float) -> float:
    """2 Re tr(x^+ y (rho + metric_reg I))."""
    return float(2 * np.vdot(x, y @ (rho + metric_reg * np.eye(len(rho)))).real)


def retract(r, k, direction, alpha: float):
    """(R + alpha W, K + i alpha/2 (W^+ R - R^+ W)), K made exactly Hermitian."""
    r_new = r + alpha * direction
    k_new = k + 0.5j * alpha * (direction.conj().T @ r - r.conj().T @ direction)
    return r_new, 0.5 * (k_new + k_new.conj().T)


class _Geometry:
    def __init__(self, rho, metric_reg: float):
        self.rho = rho
        self.reg = metric_reg
        self.weight = rho + metric_reg * np.eye(len(rho))

    def inner(self, x, y) -> float:
        return float(2 * np.vdot(x, y @ self.weight).real)

    def norm(self, x) -> float:
        sq = self.inner(x, x)
        if not (math.isfinite(sq) and sq >= 0):
            raise SolverFailure(f"tangent norm squared is {sq}")
        return math.sqrt(sq)

    def precondition(self, g, precond_reg: float):
        """P g = [(rho + precond_reg)^-1 (g (rho + metric_reg))^+]^+."""
        lhs = self.rho + precond_reg * np.eye(len(self.rho))
        return np.linalg.solve(lhs, (g @ self.weight).conj().T).conj().T


@dataclass
class _Point:
    """An evaluated point: the state (R, K), its density, energy and Riemannian gradient."""
    r: np.ndarray
    k: np.ndarray
    rho: np.ndarray
    energy: float
    gradient: np.ndarray
    stats: object = None


# ---------------------------------------------------------------------------------------------
# L-BFGS memory
# ---------------------------------------------------------------------------------------------
class _Memory:
    def __init__(self, depth: int, pairs=()):
        self.depth = depth
        self.pairs = list(pairs)[-depth:]

    def __len__(self):
        return len(self.pairs)

    def clear(self):
        self.pairs = []

    def offer(self, geo: _Geometry, step, change, gradient_norm: float) -> bool:
        """Store (step, change) if its curvature passes the cautious rule; True when stored."""
        ss, sy = geo.inner(step, step), geo.inner(step, change)
        if not (math.isfinite(ss) and math.isfinite(sy) and ss > 0 and sy / ss > gradient_norm / 1e4):
            return False
        length = math.sqrt(ss)
        self.pairs = (self.pairs + [(step / length, change / length, ss / sy)])[-self.depth:]
        return True

    def keep_curved(self, geo: _Geometry, gradient_norm: float):
        """Re-measure every pair in the current metric and drop those the cautious rule rejects."""
        kept = []
        for s, y, _ in self.pairs:
            ss, sy = geo.inner(s, s), geo.inner(s, y)
            if ss > 0 and sy / ss > gradient_norm / 1e4:
                kept.append((s, y, 1 / sy))
        self.pairs = kept

    def direction(self, geo: _Geometry, g, precond_reg: float):
        """-H g by the two-loop recursion with the preconditioner scaled by the newest pair's
        <s,y>/<y,P y>; (direction, scale), direction None when the scale is not positive."""
        alphas = []
        q = g.copy()
        for s, y, inv in self.pairs[::-1]:
            a = inv * geo.inner(s, q)
            q -= a * y
            alphas.append(a)
        s_new, y_new, _ = self.pairs[-1]
        curvature = geo.inner(y_new, geo.precondition(y_new, precond_reg))
        scale = geo.inner(s_new, y_new) / curvature if curvature > 0 else math.nan
        if not (math.isfinite(scale) and scale > 0):
            return None, scale
        h = scale * geo.precondition(q, precond_reg)
        for (s, y, inv), a in zip(self.pairs, alphas[::-1]):
            h += (a - inv * geo.inner(y, h)) * s
        return -h, float(scale)


# ---------------------------------------------------------------------------------------------
# guarded mode
# ---------------------------------------------------------------------------------------------
class _Guard:
    def __init__(self, rc: RobustnessControls, energy_atol: float):
        self.rc = rc
        self.energy_atol = energy_atol
        self.level = 0        # 0: L-BFGS; 1: memory dropped; 2: plain gradient direction
        self.streak = 0       # consecutive accepted steps the floors did not resolve
        self.e_floor = self.s_floor = 0.0

    def measure(self, point: _Point):
        """Energy and step floors at the current point."""
        self.e_floor = max(self.energy_atol, self.rc.energy_resolution_factor * EPS * max(1.0, abs(point.energy)))
        self.s_floor = self.rc.step_resolution_factor * EPS * max(1.0, float(np.linalg.norm(point.r)))

    def trial_below_floor(self, predicted: float, step_norm: float) -> bool:
        return predicted <= self.rc.predicted_decrease_factor * self.e_floor or step_norm <= self.s_floor

    def can_climb(self) -> bool:
        return self.level < self.rc.max_recoveries

    def judge_step(self, moved: float, decrease: float, predicted: float, change_norm: float, wolfe: bool):
        """(secant resolved, step resolved) of an accepted step."""
        gradient_moved = change_norm > self.rc.secant_gradient_floor
        if wolfe:   # Wolfe slopes resolve the secant even where the energy cannot
            secant = moved > self.s_floor and gradient_moved
        else:
            secant = (moved > self.s_floor and gradient_moved and decrease > self.e_floor
                      and predicted > self.rc.predicted_decrease_factor * self.e_floor)
        step = moved > self.s_floor and (decrease > self.e_floor or gradient_moved)
        return secant, step

    def after_step(self, resolved: bool) -> bool:
        """Update the level after an accepted step; True when unresolved steps have run too long."""
        if resolved:
            self.level = self.streak = 0
            return False
        self.streak += 1
        self.level = min(self.rc.max_recoveries, self.level + 1)
        return self.streak > self.rc.max_recoveries


# ---------------------------------------------------------------------------------------------
# line searches
# ---------------------------------------------------------------------------------------------
class _NoPoint(Exception):
    """A line search ends without a point (its message says why)."""


class _OutOfTime(Exception):
    pass


@dataclass
class _Found:
    point: _Point | None
    alpha: float
    trials: int
    below_floor: bool = False     # the search stopped at a trial under the guard's floors
    wolfe: dict | None = None     # Wolfe search summary, when one ran


class _Searcher:
    def __init__(self, opt: "_Optimizer", direction, slope: float, alpha0: float, alpha_max: float):
        self.opt = opt
        self.c = opt.c
        self.d = direction
        self.d_norm = float(np.linalg.norm(direction))
        self.slope = slope
        self.alpha0 = alpha0
        self.alpha_max = alpha_max
        self.failures = []

    def event(self, **fields_):
        return dict(type="trial", iteration=self.opt.iteration + 1, evaluations=self.opt.evaluations, **fields_)

    def evaluate(self, alpha: float) -> _Point:
        return self.opt.evaluate(self.d, alpha)

    # bounded Armijo backtracking: alpha0 shrink^j for j < line_search_trials
    def backtrack(self, tag):
        c, opt, guard = self.c, self.opt, self.opt.guard
        alpha = self.alpha0
        for n in range(1, c.line_search_trials + 1):
            if opt.out_of_time():
                raise _OutOfTime
            predicted = -alpha * self.slope
            if guard is not None and guard.trial_below_floor(predicted, alpha * self.d_norm):
                yield self.event(trial=n, alpha=float(alpha), accepted=False, robust_skipped_unresolved=True,
                                 expected_decrease=float(predicted), energy_resolution=guard.e_floor,
                                 nominal_step_norm=float(alpha * self.d_norm), step_resolution=guard.s_floor,
                                 recovery_mode=guard.level, **tag)
                return _Found(None, alpha, n, below_floor=True)
            bound = opt.point.energy + c.armijo_c1 * alpha * self.slope + c.energy_atol
            try:
                p = self.evaluate(alpha)
            except (SolverFailure, RuntimeError) as err:
                self.failures.append(str(err))
                yield self.event(trial=n, alpha=float(alpha), accepted=False, error=str(err), **tag)
            else:
                ok = p.energy <= bound
                yield self.event(trial=n, alpha=float(alpha), energy=p.energy, armijo_bound=float(bound),
                                 accepted=bool(ok), stats=p.stats, **tag)
                if ok:
                    return _Found(p, alpha, n)
            alpha *= c.shrink
        return _Found(None, alpha, c.line_search_trials)

    # the Hager-Zhang search of SafeguardedLineSearch, its result re-checked
    def wolfe(self):
        hz = HagerZhangLineSearch(epsilon=self.opt.guard.rc.wolfe_energy_atol)
        origin = Trial(0.0, self.opt.point.energy, float(self.slope))
        log = []

        def probe(alpha):
            if self.opt.out_of_time():
                raise _OutOfTime
            if alpha > self.alpha_max:
                raise _NoPoint(f"step {alpha:.3e} beyond the bound {self.alpha_max:.3e}")
            if alpha * self.d_norm <= self.opt.guard.s_floor:
                raise _NoPoint(f"step {alpha:.3e} under the step resolution")
            try:
                p = self.evaluate(alpha)
            except (SolverFailure, RuntimeError) as err:
                self.failures.append(str(err))
                log.append(self.event(trial=len(log) + 1, line_search="wolfe", alpha=float(alpha), accepted=False,
                                      error=str(err)))
                raise _NoPoint(str(err)) from None
            t = Trial(float(alpha), p.energy, _Geometry(p.rho, self.c.metric_reg).inner(p.gradient, self.d), p)
            exact, approx = hz.conditions(t, origin)
            log.append(self.event(trial=len(log) + 1, line_search="wolfe", alpha=t.alpha, energy=t.f, slope=t.slope,
                                  initial_slope=origin.slope, exact_wolfe=exact, approximate_wolfe=approx,
                                  accepted=False, stats=p.stats))
            return t

        t, kind, why = None, None, None
        try:
            t, _ = hz.search(probe, origin, self.alpha0, acceptfirst=True)
            kind = _wolfe_kind(t, origin, hz)
            if kind is None:
                why = f"no Wolfe condition at the returned step {t.alpha:.3e}"
        except _NoPoint as err:
            why = str(err)
        except SolverFailure as err:
            self.failures.append(str(err))
            why = str(err)
        except _OutOfTime:
            yield from log
            raise
        yield from log
        yield self.event(trial=len(log), line_search="wolfe", stage="result", alpha=t.alpha if kind else None,
                         accepted=kind is not None, wolfe_condition=kind, armijo_fallback=kind is None, reason=why)
        summary = dict(line_search="wolfe", wolfe_condition=kind, wolfe_trials=len(log), wolfe_failure=why,
                       armijo_fallback=kind is None)
        if kind is not None:
            return _Found(t.data, t.alpha, len(log), wolfe=summary)
        found = yield from self.backtrack(dict(line_search="armijo"))
        found.trials += len(log)
        found.wolfe = summary
        return found

    def run(self):
        guard = self.opt.guard
        if guard is not None and guard.rc.line_search == "wolfe":
            return (yield from self.wolfe())
        return (yield from self.backtrack({}))


def _wolfe_kind(t: Trial, origin: Trial, hz: HagerZhangLineSearch):
    """"exact", "approximate" or None: the Wolfe condition the returned trial meets."""
    if not (t.alpha > 0 and math.isfinite(t.f) and math.isfinite(t.slope)):
        return None
    curvature = t.slope >= hz.c2 * origin.slope
    if curvature and t.f - origin.f <= hz.c1 * t.alpha * origin.slope:
        return "exact"
    if curvature and t.f <= origin.f + hz.epsilon and t.slope <= (2 * hz.c1 - 1) * origin.slope:
        return "approximate"
    return None


# ---------------------------------------------------------------------------------------------
# the optimizer
# ---------------------------------------------------------------------------------------------
class _Optimizer:
    def __init__(self, point: _Point, memory: _Memory, c: OptimizationControls, guard, objective, stationary,
                 counters: dict):
        self.point = point
        self.memory = memory
        self.c = c
        self.guard = guard
        self.objective = objective
        self.stationary = stationary
        self.iteration = counters["iteration"]
        self.evaluations = counters["evaluations"]
        self.precond_reg = counters["precond_reg"]
        self.next_step = counters["next_step"]
        self.elapsed_before = counters["elapsed_seconds"]
        self.started = perf_counter()

    def geometry(self, point=None) -> _Geometry:
        return _Geometry((point or self.point).rho, self.c.metric_reg)

    def out_of_time(self) -> bool:
        return self.c.max_seconds is not None and perf_counter() - self.started >= self.c.max_seconds

    def snapshot(self) -> dict:
        p = self.point
        return dict(R=p.r, K=p.k, rho=p.rho, gradient=p.gradient, energy=p.energy, pairs=list(self.memory.pairs),
                    precond_reg=self.precond_reg, next_step=self.next_step, iteration=self.iteration,
                    evaluations=self.evaluations, elapsed_seconds=self.elapsed_before + perf_counter() - self.started,
                    gradient_norm=self.geometry().norm(p.gradient), stats=p.stats)

    def stop(self, reason: str, converged: bool = False, **extra) -> dict:
        return dict(type="stopped", reason=reason, converged=converged, state=self.snapshot(), **extra)

    def evaluate(self, direction, alpha: float) -> _Point:
        r, k = retract(self.point.r, self.point.k, direction, alpha)
        rho, _ = self.stationary(-1j * k - r.conj().T @ r / 2, r, self.c.stationary_tolerance)
        self.evaluations += 1
        energy, gradient, stats = self.objective(r, k, rho)
        energy, gradient = float(energy), np.asarray(gradient, dtype=np.complex128)
        if not (math.isfinite(energy) and gradient.shape == r.shape and np.isfinite(gradient).all()):
            raise SolverFailure("a trial returned a nonfinite or misshapen energy or gradient")
        return _Point(r, k, np.asarray(rho, dtype=np.complex128), energy, gradient, stats)

    def direction(self, geo: _Geometry, gnorm: float):
        """(direction, slope, scale): by the guard's level, the memory, or the scaled preconditioned
        gradient; whatever is not a descent direction falls back to the last."""
        g = self.point.gradient
        d, scale = None, None
        if self.guard is not None and self.guard.level == 2:
            d = -(self.c.initial_direction_norm / gnorm) * g
        elif len(self.memory):
            d, scale = self.memory.direction(geo, g, self.precond_reg)
        slope = geo.inner(g, d) if d is not None else math.nan
        if not (slope < 0):
            self.memory.clear()
            pg = geo.precondition(g, self.precond_reg)
            size = geo.norm(pg)
            if size == 0:
                raise SolverFailure("the preconditioned gradient vanished before convergence")
            d = -(self.c.initial_direction_norm / size) * pg
            slope = geo.inner(g, d)
            if not (slope < 0):
                raise SolverFailure("the preconditioned gradient gives no descent")
        return d, slope, scale

    def accept(self, found: _Found, direction, slope: float):
        """Move to the found point, update the memory, the preconditioner and the guard; returns the
        accepted event and whether the guard ends the run."""
        old, new = self.point, found.point
        geo_new = self.geometry(new)
        step, change = found.alpha * direction, new.gradient - old.gradient
        gnorm = geo_new.norm(new.gradient)
        info, store, resolved, stop = {}, True, True, False
        if self.guard is not None:
            moved = float(np.linalg.norm(new.r - old.r))
            decrease = old.energy - new.energy
            predicted = -found.alpha * slope
            change_norm = geo_new.norm(change)
            store, resolved = self.guard.judge_step(moved, decrease, predicted, change_norm, found.wolfe is not None)
            info = dict(robust_secant_resolved=bool(store), actual_step_norm=moved, step_resolution=self.guard.s_floor,
                        gradient_change_norm=change_norm, actual_energy_decrease=float(decrease),
                        expected_decrease=float(predicted), energy_resolution=self.guard.e_floor)
        stored = store and self.memory.offer(geo_new, step, change, gnorm)
        self.point = new
        self.iteration += 1
        self.next_step = min(1.0, 2 * found.alpha) if self.c.warm_start_line_search else self.c.initial_step
        self.precond_reg = max(10 * EPS, self.c.precond_prefactor * min(gnorm * gnorm, 1.0))
        if self.guard is not None:
            stop = self.guard.after_step(resolved)
            if not resolved:
                self.memory.clear()
                self.next_step = self.c.initial_step
            info.update(robust_unresolved_accepted=not resolved, recovery_mode=self.guard.level)
        return info, gnorm, bool(stored), stop

    def events(self):
        c = self.c
        while True:
            geo = self.geometry()
            gnorm = geo.norm(self.point.gradient)
            if gnorm <= c.gradient_tol:
                yield self.stop("converged: gradient norm within gradient_tol", converged=True)
                return
            if self.iteration >= c.max_iterations:
                yield self.stop("iteration cap max_iterations reached")
                return
            if self.out_of_time():
                yield self.stop("time budget max_seconds spent")
                return
            began = perf_counter()
            if self.guard is not None:
                self.guard.measure(self.point)
            if c.refresh_curvatures:
                self.memory.keep_curved(geo, gnorm)
            direction, slope, scale = self.direction(geo, gnorm)
            reach = c.max_relative_step * max(1.0, float(np.linalg.norm(self.point.r)))
            alpha_max = reach / max(float(np.linalg.norm(direction)), EPS)
            search = _Searcher(self, direction, slope, min(self.next_step, alpha_max), alpha_max)
            try:
                found = yield from search.run()
            except _OutOfTime:
                yield self.stop("time budget max_seconds spent during a line search",
                                numerical_failures=search.failures)
                return
            if found.point is None:
                if self.guard is None:
                    yield self.stop("line search failed; the last accepted point is kept",
                                    numerical_failures=search.failures)
                    return
                if self.guard.can_climb():
                    self.guard.level += 1
                    self.memory.clear()
                    self.next_step = c.initial_step
                    yield dict(type="trial", iteration=self.iteration + 1, trial=found.trials,
                               alpha=float(self.next_step), accepted=False, robust_restart=True,
                               reason="trial under the resolution floors" if found.below_floor
                               else "line search failed", recovery_mode=self.guard.level, history_size=0,
                               evaluations=self.evaluations)
                    continue
                if found.below_floor:
                    yield self.stop("stagnation: trials fall under the resolution floors at the last recovery "
                                    "level", numerical_failures=search.failures, recovery_mode=self.guard.level)
                else:
                    yield self.stop("line search failed at the last recovery level; the last accepted point is "
                                    "kept", numerical_failures=search.failures, recovery_mode=self.guard.level)
                return
            info, gnorm_new, stored, stop = self.accept(found, direction, slope)
            wolfe = found.wolfe or {}
            yield dict(type="accepted", state=self.snapshot(), iteration=self.iteration, energy=self.point.energy,
                       gradient_norm=gnorm_new, alpha=float(found.alpha), trials=found.trials,
                       seconds=perf_counter() - began, stats=self.point.stats, gamma=scale, cautious_update=stored,
                       history_size=len(self.memory), evaluations=self.evaluations, **info, **wolfe)
            if stop and gnorm_new > c.gradient_tol:
                yield self.stop("stagnation: accepted steps stay under the resolution floors",
                                unresolved_acceptances=self.guard.streak)
                return


def _start(r, k, rho, energy, gradient, c: OptimizationControls, resume_state):
    """The starting point, memory and counters from the arguments or a snapshot, validated."""
    snap = resume_state or {}
    as_c = lambda key, default: np.array(snap.get(key, default), dtype=np.complex128)  # noqa: E731
    point = _Point(as_c("R", r), as_c("K", k), as_c("rho", rho), float(snap.get("energy", energy)),
                   as_c("gradient", gradient), snap.get("stats"))
    arrays = (point.k, point.rho, point.gradient)
    if (point.r.ndim != 2 or point.r.shape[0] != point.r.shape[1] or any(a.shape != point.r.shape for a in arrays)
            or not math.isfinite(point.energy) or not all(np.isfinite(a).all() for a in (point.r,) + arrays)):
        raise ValueError("invalid starting point")
    pairs = []
    for s, y, inv in snap.get("pairs", []):
        s, y, inv = np.array(s, dtype=np.complex128), np.array(y, dtype=np.complex128), float(inv)
        if s.shape != point.r.shape or y.shape != point.r.shape or not _positive(inv) \
                or not (np.isfinite(s).all() and np.isfinite(y).all()):
            raise ValueError("invalid L-BFGS pair in the resumed state")
        pairs.append((s, y, inv))
    counters = dict(iteration=int(snap.get("iteration", 0)), evaluations=int(snap.get("evaluations", 0)),
                    precond_reg=float(snap.get("precond_reg", c.initial_precond_reg)),
                    next_step=float(snap.get("next_step", c.initial_step)),
                    elapsed_seconds=float(snap.get("elapsed_seconds", 0.0)))
    if (counters["iteration"] < 0 or counters["evaluations"] < 0 or not _positive(counters["precond_reg"])
            or not _positive(counters["next_step"])
            or not (math.isfinite(counters["elapsed_seconds"]) and counters["elapsed_seconds"] >= 0)):
        raise ValueError("invalid counters in the resumed state")
    return point, _Memory(c.history, pairs), counters


def optimize(r, k, rho, energy, gradient, objective, controls=None, resume_state=None, *, robustness=None,
             stationary=None):
    """Yield trial / accepted / stopped events from the evaluated point (r, k, rho, energy, gradient),
    or from `resume_state` (a snapshot of an earlier run, which then replaces them)."""
    c = check_controls(controls)
    point, memory, counters = _start(r, k, rho, energy, gradient, c, resume_state)
    guard = _Guard(robustness, c.energy_atol) if robustness is not None else None
    opt = _Optimizer(point, memory, c, guard, objective, stationary or gpu_stationary(), counters)
    yield from opt.events()


# ---------------------------------------------------------------------------------------------
# evaluators on the GPU
# ---------------------------------------------------------------------------------------------
class _LastState:
    """The State of the last stationary solve, which the objective reuses for the same trial."""
    state = None


def gpu_stationary(device: int = 0):
    """stationary(q, r, tolerance) -> (rho, None): the trial's stationary density solved on the GPU
    (State, src/state.jl:40-53), with K = i (Q + R^+R/2)."""
    from .state import State

    def solve(q, r, tolerance):
        k = 1j * (q + r.conj().T @ r / 2)
        s = State(r, 0.5 * (k + k.conj().T), steady_tol=tolerance, device=device)
        _LastState.state = s
        return s.rho, None
    return solve


def gpu_objective(coupling: float, *, integrator=None, device: int = 0, precision: str = "fp64"):
    """objective(r, k, rho) -> (energy, gradient, stats): fg_phi4 at this coupling on the GPU, on the
    State of the last stationary solve when it is this trial, else on a new one with rho checked."""
    from .observables import fg_phi4
    from .state import State

    def evaluate(r, k, rho):
        s = _LastState.state
        if s is None or not (np.array_equal(s.R, r) and np.array_equal(s.K, k)):
            s = State(r, k, rho=rho, device=device)
        if integrator is not None:
            s.integrator_param = integrator
        return fg_phi4(s, float(coupling), return_stats=True, precision=precision)
    return evaluate


# ---------------------------------------------------------------------------------------------
# snapshots on disk
# ---------------------------------------------------------------------------------------------
_SCHEMA = "rcmps_gpu resume state 1"
_SCALARS = ("energy", "precond_reg", "next_step", "iteration", "evaluations", "elapsed_seconds")


def save_resume_state(path, state) -> Path:
    """Write a snapshot to one .npz file (arrays, the memory pairs, the counters as JSON bytes),
    atomically through a temporary file; no pickles."""
    path = Path(path)
    pairs = state.get("pairs", [])
    d = np.asarray(state["R"]).shape[0]
    stack = lambda i: np.array([p[i] for p in pairs], dtype=np.complex128).reshape(len(pairs), d, d)  # noqa: E731
    meta = {key: state[key] for key in _SCALARS if key in state}
    meta["schema"] = _SCHEMA
    content = {key: np.asarray(state[name], dtype=np.complex128)
               for key, name in (("R", "R"), ("K", "K"), ("rho", "rho"), ("gradient", "gradient"))}
    content.update(pair_s=stack(0), pair_y=stack(1), pair_inv=np.array([p[2] for p in pairs], dtype=np.float64),
                   meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    path.parent.mkdir(parents=True, exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=path.parent, suffix=".npz.tmp")
    try:
        with os.fdopen(fd, "wb") as f:
            np.savez(f, **content)
        os.replace(tmp, path)
    finally:
        if os.path.exists(tmp):
            os.remove(tmp)
    return path


def load_resume_state(path) -> dict:
    """A snapshot written by save_resume_state, ready for optimize(resume_state=...)."""
    with np.load(Path(path), allow_pickle=False) as z:
        meta = json.loads(bytes(z["meta"]).decode())
        if meta.get("schema") != _SCHEMA:
            raise ValueError("not an rcmps_gpu resume state")
        state = {key: z[key].copy() for key in ("R", "K", "rho", "gradient")}
        state["pairs"] = [(z["pair_s"][i].copy(), z["pair_y"][i].copy(), float(z["pair_inv"][i]))
                          for i in range(len(z["pair_inv"]))]
    state.update({key: meta[key] for key in _SCALARS if key in meta})
    return state
