"""Line searches of optimize_state (src/optimization.jl:24-106).

HagerZhangLineSearch follows the approximate-Wolfe line search of Hager and Zhang (ACM TOMS 32,
2006, "Algorithm 851: CG_DESCENT") as the pinned OptimKit 0.5.1 runs it inside RcmpsCore's
SafeguardedLineSearch:

- the first trial is accepted when it satisfies either Wolfe condition;
- otherwise a bracket is grown by factor rho (no checks at the expansion points), or bisected
  from the origin when a trial is uphill with a negative slope;
- then secant2 steps, with a bisection whenever the bracket shrinks by less than gamma;
- every point an update produces is checked;
- the search stops after max_iterations secant rounds or max_evaluations trials and then returns
  its last lower bracket end.

SafeguardedLineSearch bounds the step and re-checks the point returned by the search (exact
Wolfe, or approximate Wolfe with the energy within energy_atol). It falls back to Armijo
backtracking on a SolverFailure or an invalid return.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from ._lib import SolverFailure


@dataclass(frozen=True)
class Trial:
    """A point on the search line: step, energy, slope d/dalpha, and what a caller keeps."""
    alpha: float
    f: float
    slope: float
    data: object = None


EPS = float(np.finfo(float).eps)


@dataclass
class HagerZhangLineSearch:
    c1: float = 0.1
    c2: float = 0.9
    epsilon: float = 1e-6       # energy tolerance of the approximate Wolfe condition
    theta: float = 0.5
    gamma: float = 2 / 3
    rho: float = 5.0
    max_iterations: int = 10
    max_evaluations: int = 20

    def conditions(self, t: Trial, origin: Trial):
        """(exact, approximate) Wolfe conditions of trial t."""
        exact = t.f <= origin.f + self.c1 * t.alpha * origin.slope and t.slope > self.c2 * origin.slope
        approx = (t.f <= origin.f + self.epsilon and
                  (2 * self.c1 - 1) * origin.slope >= t.slope >= self.c2 * origin.slope)
        return bool(exact), bool(approx)

    def __call__(self, fg, x, p, fg0, initialguess=1.0, *, retract, inner, acceptfirst=False):
        """OptimKit's HagerZhangLineSearch used directly as a line search (ls_lenient,
        src/optimization.jl:106): the search's final trial is returned without a re-check."""
        f0, g0 = fg0
        slope0 = inner(x, g0, p)
        if not (math.isfinite(slope0) and slope0 < 0):
            raise SolverFailure("line search direction is not a descent direction")
        calls = [0]

        def evaluate(alpha):
            y, v = retract(x, p, alpha)
            calls[0] += 1
            f, g = fg(y)
            return Trial(alpha, f, inner(y, g, v), (y, g, v))

        t, _ = self.search(evaluate, Trial(0.0, f0, slope0, (x, g0, p)), initialguess, acceptfirst)
        y, g, v = t.data
        return y, t.f, g, v, t.alpha, calls[0]

    def search(self, evaluate, origin: Trial, alpha: float, acceptfirst: bool = True):
        """evaluate(step) -> Trial.  Returns (final trial, list of trial steps).  With acceptfirst
        (what OptimKit's L-BFGS passes) a first trial meeting either Wolfe condition is returned
        at once; otherwise it only starts the bracket."""
        return _Search(self, evaluate, origin).run(alpha, acceptfirst)


class _Search:
    def __init__(self, hz: HagerZhangLineSearch, evaluate, origin: Trial):
        self.hz, self.evaluate, self.origin = hz, evaluate, origin
        self.fmax = origin.f + hz.epsilon
        self.steps = []

    def take(self, alpha) -> Trial:
        self.steps.append(alpha)
        return self.evaluate(alpha)

    def ok(self, t: Trial) -> bool:
        return any(self.hz.conditions(t, self.origin))

    @staticmethod
    def secant(a: Trial, b: Trial) -> float:
        if b.slope == a.slope:
            return math.nan        # outside every bracket: the update leaves it unchanged
        return (a.alpha * b.slope - b.alpha * a.slope) / (b.slope - a.slope)

    def bisect(self, a: Trial, b: Trial):
        """U3a-c: a has slope < 0 and energy <= fmax, b slope < 0 and energy > fmax.
        Returns (a, b, evaluations)."""
        n = 0
        while b.alpha - a.alpha > EPS ** 0.75 and n < self.hz.max_evaluations:
            d = self.take((1 - self.hz.theta) * a.alpha + self.hz.theta * b.alpha)
            n += 1
            if self.ok(d):
                return d, d, n
            if d.slope >= 0:
                return a, d, n
            if d.f <= self.fmax:
                a = d
            else:
                b = d
        return a, b, n

    def update(self, a: Trial, b: Trial, alpha: float):
        """U0-U3 with a trial at alpha.  Returns (a, b, evaluations)."""
        if not (a.alpha < alpha < b.alpha):
            return a, b, 0
        c = self.take(alpha)
        if not math.isfinite(c.f):
            raise SolverFailure(f"line search: nonfinite energy at step {alpha}")
        if self.ok(c):
            return c, c, 1
        if c.slope >= 0:
            return a, c, 1
        if c.f <= self.fmax:
            return c, b, 1
        a2, b2, n = self.bisect(a, c)
        return a2, b2, n + 1

    def bracket(self, c: Trial):
        """B0-B3 from the first trial c.  Returns (a, b, evaluations after the first)."""
        hz, a, n = self.hz, self.origin, 0
        alpha = c.alpha
        while True:
            while not (math.isfinite(c.f) and math.isfinite(c.slope)) and n < hz.max_evaluations:
                alpha = (a.alpha + alpha) / 2         # step back towards the last good point
                c = self.take(alpha)
                n += 1
            if not (math.isfinite(c.f) and math.isfinite(c.slope)):
                return a, a, n
            if c.slope >= 0:
                return a, c, n
            if c.f > self.fmax:
                a2, b2, m = self.bisect(self.origin, c)   # from the origin, not from the last a
                return a2, b2, n + m
            a = c
            if n >= hz.max_evaluations:
                return a, a, n
            alpha *= hz.rho
            c = self.take(alpha)
            n += 1

    def run(self, alpha: float, acceptfirst: bool):
        hz = self.hz
        c = self.take(alpha)
        if acceptfirst and self.ok(c):
            return c, self.steps
        a, b, n = self.bracket(c)
        evaluations = n + 1
        rounds = 1
        while a.alpha != b.alpha and b.alpha - a.alpha >= EPS:
            if rounds >= hz.max_iterations or evaluations >= hz.max_evaluations:
                break
            width = b.alpha - a.alpha
            s = self.secant(a, b)
            A, B, m = self.update(a, b, s)
            evaluations += m
            if A.alpha == B.alpha:
                return A, self.steps
            if s == B.alpha:
                a, b, m = self.update(A, B, self.secant(b, B))
            elif s == A.alpha:
                a, b, m = self.update(A, B, self.secant(a, A))
            else:
                a, b, m = A, B, 0
            evaluations += m
            if a.alpha != b.alpha and b.alpha - a.alpha > hz.gamma * width:
                a, b, m = self.update(a, b, (a.alpha + b.alpha) / 2)
                evaluations += m
            rounds += 1
        return a, self.steps


@dataclass
class BackwardLineSearch:
    """Bounded Armijo backtracking; a failed search never returns an uphill step
    (src/optimization.jl:24-53).  Trials whose retraction or evaluation raise SolverFailure are
    rejected and the step is shrunk."""
    c1: float = 1e-4
    shrink: float = 0.5
    maxiter: int = 30

    def __call__(self, fg, x, p, fg0, initialguess=1.0, *, retract, inner, acceptfirst=False):
        f0, g0 = fg0
        slope = inner(x, g0, p)
        if not slope < 0:
            raise SolverFailure("line search direction is not a descent direction")
        alpha = initialguess
        failures, last = 0, ""
        for k in range(1, self.maxiter + 1):
            try:
                y, v = retract(x, p, alpha)
                f, g = fg(y)
                if math.isfinite(f) and f <= f0 + self.c1 * alpha * slope:
                    return y, f, g, v, alpha, k
            except SolverFailure as e:
                failures += 1
                last = str(e)
            alpha *= self.shrink
        detail = f"; {failures} numerical trial failures; last cause: {last}" if failures else ""
        raise SolverFailure(f"Armijo line search exhausted {self.maxiter} steps{detail}")


class SafeguardedLineSearch:
    """Hager-Zhang search with a bounded step and a checked Armijo recovery
    (src/optimization.jl:54-103).

    Trials beyond alpha_max = max_relative_step max(1, |R|_F) / |p|_F raise SolverFailure.  The
    point the Wolfe search returns (possibly its last trial after exhausting its budget) is
    re-checked: exact Wolfe (f - f0 <= c1 alpha slope0 and slope >= c2 slope0) or approximate Wolfe
    (f <= f0 + energy_atol and c2 slope0 <= slope <= (2 c1 - 1) slope0), with finite f and slope and
    alpha > 0.  A SolverFailure anywhere in the Wolfe search, or a point that fails the re-check,
    hands over to BackwardLineSearch from the same initial guess.  The returned evaluation count
    includes the abandoned Wolfe trials."""

    def __init__(self, max_relative_step: float = 0.5, energy_atol: float = 1e-10):
        if not (math.isfinite(energy_atol) and energy_atol >= 0):
            raise ValueError("invalid energy tolerance")
        self.wolfe = HagerZhangLineSearch(epsilon=energy_atol)
        self.max_relative_step = float(max_relative_step)
        self.backtracking = BackwardLineSearch()
        self.energy_atol = float(energy_atol)
        self.last_used = None          # "wolfe" or "backtracking": which search returned the step

    def __call__(self, fg, x, p, fg0, initialguess=1.0, *, retract, inner, acceptfirst=True):
        if not self.max_relative_step > 0:
            raise ValueError("step bound must be positive")
        f0, g0 = fg0
        slope0 = inner(x, g0, p)
        if not (math.isfinite(slope0) and slope0 < 0):
            raise SolverFailure("line search direction is not a descent direction")
        alpha_max = self.max_relative_step * max(1.0, float(np.linalg.norm(x.R))) / max(float(np.linalg.norm(p)), EPS)
        calls = [0]

        def counted(y):
            calls[0] += 1
            return fg(y)

        def bounded(x_, p_, alpha):
            if abs(alpha) > alpha_max:
                raise SolverFailure("line search trial exceeds step bound")
            return retract(x_, p_, alpha)

        def evaluate(alpha):
            y, v = bounded(x, p, alpha)
            f, g = counted(y)
            return Trial(alpha, f, inner(y, g, v), (y, g, v))

        alpha0 = min(initialguess, alpha_max)
        try:
            t, _ = self.wolfe.search(evaluate, Trial(0.0, f0, slope0, (x, g0, p)), alpha0, acceptfirst)
            c1, c2 = self.wolfe.c1, self.wolfe.c2
            exact = t.f - f0 <= c1 * t.alpha * slope0 and t.slope >= c2 * slope0
            approx = t.f <= f0 + self.energy_atol and c2 * slope0 <= t.slope <= (2 * c1 - 1) * slope0
            if not (math.isfinite(t.f) and math.isfinite(t.slope) and t.alpha > 0 and (exact or approx)):
                raise SolverFailure("Wolfe search returned a trial without valid Wolfe conditions")
            y, g, v = t.data
            self.last_used = "wolfe"
            return y, t.f, g, v, t.alpha, calls[0]
        except SolverFailure:
            y, f, g, v, alpha, _ = self.backtracking(counted, x, p, fg0, alpha0, retract=bounded, inner=inner)
            self.last_used = "backtracking"
            return y, f, g, v, alpha, calls[0]
