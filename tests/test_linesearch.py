"""Hager-Zhang line search against OptimKit 0.5.1, step for step.

The golden table (tests/golden/hager_zhang_trials.json, from the reference's own recorded
values) lists, for ten one-dimensional functions phi(alpha) with their slopes, every trial step
OptimKit's HagerZhangLineSearch(epsilon=1e-10) takes with acceptfirst=true (as its L-BFGS calls
it) and the step it returns.  The functions cover: acceptance of the first trial, bracket growth
by rho, secant steps, approximate Wolfe acceptance under a constant energy offset, bisection
towards an energy wall, the evaluation budget on an unbounded linear function, and the secant2
bisection safeguard.  Reproducing every trial sequence exactly pins the order and arithmetic of
the search.  Powers are written as products, as in the recorded functions.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from rcmps_gpu.linesearch import HagerZhangLineSearch, Trial, SafeguardedLineSearch
from rcmps_gpu._lib import SolverFailure

GOLDEN = json.loads((Path(__file__).parent / "golden" / "hager_zhang_trials.json").read_text())["cases"]


def _wall(x, lower, upper, height):
    return height if lower < x < upper else 0.0


def _bump(a):
    z = (a - 0.3) / 0.15
    return math.exp(-z * z)


FUNCTIONS = {
    "quadratic_min_3": (lambda a: ((a - 3) * (a - 3), 2 * (a - 3)), 1.0),
    "quadratic_min_20": (lambda a: ((a - 20) * (a - 20), 2 * (a - 20)), 1.0),
    "exponential": (lambda a: (math.exp(a) - 5 * a, math.exp(a) - 5), 10.0),
    "offset_energy": (lambda a: (1e-12 * (a - 1) * (a - 1) + (0.0 if a == 0 else 5e-12), 2e-12 * (a - 1)), 3.0),
    "energy_wall": (lambda a: (-a + (1e3 if a > 0.5 else 0.0), -1.0), 1.0),
    "cosine": (lambda a: (math.cos(0.3 + a) + 0.1 * (0.3 + a), -math.sin(0.3 + a) + 0.1), 1.0),
    "linear": (lambda a: (-a, -1.0), 1e-3),
    "steep_wall": (lambda a: (-a + 50 * max(0.0, a - 0.7) * max(0.0, a - 0.7), -1 + 100 * max(0.0, a - 0.7)), 8.0),
    "gaussian_bump": (lambda a: (-3 * a + a * a * a / 3 + 3 * _bump(a),
                                 -3 + a * a - 2 * 3 * (a - 0.3) / (0.15 * 0.15) * _bump(a)), 4.0),
    "cubic_wall": (lambda a: (-1.5 * a + a * a * a / 3 + _wall(a, 0.2, 0.5, 3.0), -1.5 + a * a), 4.0),
}


def run(name, **kw):
    phi, alpha0 = FUNCTIONS[name]

    def evaluate(a):
        f, d = phi(a)
        return Trial(a, f, d)

    return HagerZhangLineSearch(epsilon=1e-10, **kw).search(evaluate, evaluate(0.0), alpha0)


def test_golden_covers_every_function():
    assert set(GOLDEN) == set(FUNCTIONS)


@pytest.mark.parametrize("name", sorted(FUNCTIONS))
def test_trial_sequence_matches_optimkit(name):
    point, steps = run(name)
    assert steps == GOLDEN[name]["trials"]
    assert point.alpha == GOLDEN[name]["step"]


def test_evaluation_budget():
    # linear: the bracket never closes; the search stops after max_evaluations trials
    _, steps = run("linear")
    assert len(steps) == 21
    _, steps = run("linear", max_evaluations=5)
    assert len(steps) == 6


def test_acceptfirst_false_still_brackets():
    phi, _ = FUNCTIONS["quadratic_min_3"]
    ev = lambda a: Trial(a, *phi(a))
    point, steps = HagerZhangLineSearch(epsilon=1e-10).search(ev, ev(0.0), 1.0, acceptfirst=False)
    assert len(steps) > 1 and steps[0] == 1.0
    f, d = phi(point.alpha)
    assert f <= 9.0 + 0.1 * point.alpha * -6.0 and d >= 0.9 * -6.0


def test_nonfinite_update_raises():
    # a secant trial that lands on a nonfinite energy is a SolverFailure, not a step
    def evaluate(a):
        if 0.5 < a < 2.0:
            return Trial(a, math.nan, math.nan)
        return Trial(a, (a - 1) * (a - 1) + (1 if a >= 2 else 0), 2 * (a - 1))
    with pytest.raises(SolverFailure):
        HagerZhangLineSearch(epsilon=1e-10).search(evaluate, evaluate(0.0), 3.0)


class _Line:
    """A one-dimensional 'manifold point' for SafeguardedLineSearch: x.R is the position."""
    def __init__(self, r):
        self.R = np.array([[r]], dtype=float)


def _line_problem(phi):
    def retract(x, p, alpha):
        return _Line(float(x.R[0, 0] + alpha * p[0, 0])), p

    def inner(x, g, v):
        return float(g[0, 0] * v[0, 0])

    def fg(x):
        f, d = phi(float(x.R[0, 0]))
        return f, np.array([[d]])
    return fg, retract, inner


def test_safeguarded_takes_the_wolfe_step():
    phi = lambda a: ((a - 0.3) * (a - 0.3), 2 * (a - 0.3))
    fg, retract, inner = _line_problem(phi)
    x = _Line(0.0)
    ls = SafeguardedLineSearch()
    y, f, g, v, alpha, n = ls(fg, x, np.array([[1.0]]), fg(x), 1.0, retract=retract, inner=inner)
    assert ls.last_used == "wolfe"
    # alpha_max = 0.5 max(1, 0) / 1 = 0.5 bounds the first trial, which meets the (weak) Wolfe
    # conditions: 0.04 <= 0.09 - 0.1 * 0.5 * 0.6 and slope 0.4 > 0.9 * -0.6
    assert alpha == 0.5 and n == 1
    assert y.R[0, 0] == alpha and f == phi(alpha)[0]
    # a short first trial (slope -0.56 < 0.9 * -0.6) is rejected: the bracket grows to
    # (0.1, 0.5) and its secant lands on the minimum
    y, f, g, v, alpha, n = ls(fg, x, np.array([[1.0]]), fg(x), 0.02, retract=retract, inner=inner)
    assert ls.last_used == "wolfe" and n == 4
    assert abs(alpha - 0.3) < 1e-12 and abs(y.R[0, 0] - 0.3) < 1e-12


def test_safeguarded_recovers_with_backtracking():
    # linear descent never meets the curvature condition: the Wolfe search walks into the step
    # bound (SolverFailure) and Armijo backtracking from the bounded initial guess takes over
    phi = lambda a: (-a, -1.0)
    fg, retract, inner = _line_problem(phi)
    x = _Line(0.0)
    ls = SafeguardedLineSearch()
    y, f, g, v, alpha, n = ls(fg, x, np.array([[1.0]]), fg(x), 0.1, retract=retract, inner=inner)
    assert ls.last_used == "backtracking"
    assert alpha == 0.1 and f == -0.1
    assert n == 3          # 0.1 and 0.5 by the Wolfe search (2.5 is refused before evaluating), 0.1 again


def test_safeguarded_rejects_ascent_and_bad_controls():
    fg, retract, inner = _line_problem(lambda a: (a * a, 2 * a))
    x = _Line(1.0)
    with pytest.raises(SolverFailure):
        SafeguardedLineSearch()(fg, x, np.array([[1.0]]), fg(x), retract=retract, inner=inner)
    with pytest.raises(ValueError):
        SafeguardedLineSearch(energy_atol=-1.0)
    with pytest.raises(ValueError):
        SafeguardedLineSearch(max_relative_step=0.0)(fg, x, np.array([[-1.0]]), fg(x), retract=retract, inner=inner)


def test_wolfe_acceptance_at_the_noise_floor():
    # test/runtests.jl:5-20: an accurate slope with an energy offset (1e-11) larger than the true
    # decrease is accepted by the approximate Wolfe condition within energy_atol; with
    # energy_atol = 0 neither search accepts it, and an ascent direction is refused
    class Point:
        def __init__(self, position):
            self.R = np.ones((1, 1))
            self.position = position

    def retract(x, p, alpha):
        return Point(x.position + alpha * p), p

    def inner(x, g, p):
        return g * p

    def fg(x):
        z = x.position - 1e-6
        return z * z / 2 + (0.0 if x.position == 0 else 1e-11), z

    x = Point(0.0)
    ls = SafeguardedLineSearch(energy_atol=1e-10)
    y, f, g, _, alpha, _ = ls(fg, x, 1e-6, fg(x), retract=retract, inner=inner)
    assert y.position == pytest.approx(1e-6, rel=1e-12)
    assert g == 0
    assert fg(x)[0] < f <= fg(x)[0] + ls.energy_atol
    assert alpha > 0
    with pytest.raises(SolverFailure):
        SafeguardedLineSearch(energy_atol=0)(fg, x, 1e-6, fg(x), retract=retract, inner=inner)
    with pytest.raises(SolverFailure):
        ls(fg, x, -1e-6, fg(x), retract=retract, inner=inner)
