"""The resumable L-BFGS controller (rcmps_gpu.controller) on the CPU.

Objectives: the oracle's fg_phi4 at D = 2 (a real energy with its Riemannian gradient, stationary
state from the oracle), and a quadratic E = |R - R0|^2 with a fixed density, whose Riemannian
gradient is (R - R0)(rho + reg)^-1 in closed form, optionally with a deterministic energy noise.

Checks: control validation; Armijo runs lower the energy at every accepted step and keep the
normalised pairs; a run stopped and resumed from its saved snapshot reproduces the uninterrupted
run bit for bit; on a flat (quartic) minimum whose energy carries a noise of 1e-9, the Armijo
controller with resolution guards stops at numerical stagnation above 100 times the target while the
Hager-Zhang Wolfe search (energy tolerance 1e-8), whose slopes are exact, reaches |g| <= 1e-10
(the behaviour gpu/README.md:500-505 states for the reference controller);
the time budget and iteration cap stop a run with a resumable state.
"""
import hashlib
import math
import sys

import numpy as np
import pytest

import oracle
from rcmps_gpu.controller import (OptimizationControls, RobustnessControls, check_controls, load_resume_state,
                                  metric, optimize, save_resume_state)

REG = 1e-3


def oracle_problem(seed=1, d=2, g=2.0):
    rng = np.random.default_rng(seed)
    x = 0.3 * (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d)))
    k = (x + x.conj().T) / 2
    r = 0.3 * (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d)))

    def stationary(q, rr, tol):
        return oracle.steady_state(q, rr, tol=tol), None

    def objective(rr, kk, rho):
        q = -1j * kk - rr.conj().T @ rr / 2
        out = oracle.fg_phi4(rr, q, rho, g, n_quad=16, max_point=12.0, atol=1e-10, rtol=1e-10, metric_reg=REG)
        return out["energy"], out["gradient"], None

    rho, _ = stationary(-1j * k - r.conj().T @ r / 2, r, 1e-13)
    e, grad, _ = objective(r, k, rho)
    return r, k, rho, e, grad, objective, stationary


def quadratic_problem(noise=0.0, d=3, seed=4, power=1):
    """E = s^power with s = |R - R0|^2 and a fixed density; the Riemannian gradient of
    dE[W] = power s^(power-1) 2 Re tr((R - R0)^+ W) in the metric 2 Re tr(G^+ W (rho + reg)) is
    power s^(power-1) (R - R0)(rho + reg)^-1.  power = 2 makes the minimum flat (quartic)."""
    rng = np.random.default_rng(seed)
    r0 = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    rho = np.eye(d) / d
    m_inv = np.linalg.inv(rho + REG * np.eye(d))

    def stationary(q, rr, tol):
        return rho, None

    def objective(rr, kk, rh):
        sq = float(np.sum(np.abs(rr - r0) ** 2))
        e = sq ** power
        if noise:
            h = hashlib.sha256(np.ascontiguousarray(rr).tobytes()).digest()
            e += noise * (int.from_bytes(h[:8], "little") / 2 ** 64 - 0.5)
        return e, power * sq ** (power - 1) * (rr - r0) @ m_inv, None

    r = np.zeros((d, d), complex)
    k = np.zeros((d, d), complex)
    e, g, _ = objective(r, k, rho)
    return r, k, rho, e, g, objective, stationary


def run(events):
    out = list(events)
    assert out[-1]["type"] == "stopped"
    return out


def test_controls_are_checked():
    c = check_controls({"iterations": 7, "gradtol": 1e-6, "lbfgs_history": 5, "unknown": 1})
    assert (c.max_iterations, c.gradient_tol, c.history) == (7, 1e-6, 5)
    for bad in ({"history": 0}, {"gradient_tol": 0.0}, {"armijo_c1": 1.0}, {"shrink": 1.5},
                {"max_seconds": -1.0}, {"energy_atol": -1e-3}, {"line_search_trials": 0}):
        with pytest.raises(ValueError):
            check_controls(bad)
    RobustnessControls(line_search="wolfe", wolfe_energy_atol=1e-10)
    for bad in (dict(line_search="wolfe"), dict(wolfe_energy_atol=1e-10), dict(max_recoveries=3),
                dict(line_search="bisect"), dict(energy_resolution_factor=0.0)):
        with pytest.raises(ValueError):
            RobustnessControls(**bad)


def test_armijo_run_lowers_the_energy():
    r, k, rho, e, g, objective, stationary = oracle_problem()
    c = OptimizationControls(max_iterations=12, gradient_tol=1e-9, history=4)
    events = run(optimize(r, k, rho, e, g, objective, c, stationary=stationary))
    accepted = [ev for ev in events if ev["type"] == "accepted"]
    assert len(accepted) >= 8
    energies = [e] + [ev["energy"] for ev in accepted]
    assert all(b <= a for a, b in zip(energies, energies[1:]))
    assert accepted[-1]["gradient_norm"] < 0.1 * math.sqrt(metric(rho, g, g, REG))
    last = accepted[-1]["state"]
    assert len(last["pairs"]) <= 4
    assert all(inv > 0 for _, _, inv in last["pairs"])
    # pairs are normalised in the metric of the state where they were taken: the newest is this one
    s_new = last["pairs"][-1][0]
    assert abs(metric(last["rho"], s_new, s_new, REG) - 1.0) < 1e-12


def test_resume_reproduces_the_uninterrupted_run(tmp_path):
    r, k, rho, e, g, objective, stationary = oracle_problem(seed=3)
    full = run(optimize(r, k, rho, e, g, objective, OptimizationControls(max_iterations=10, gradient_tol=1e-12,
                                                                         history=3), stationary=stationary))
    first = run(optimize(r, k, rho, e, g, objective, OptimizationControls(max_iterations=4, gradient_tol=1e-12,
                                                                          history=3), stationary=stationary))
    assert first[-1]["reason"] == "iteration cap max_iterations reached"
    path = save_resume_state(tmp_path / "run" / "state.npz", first[-1]["state"])
    resumed_state = load_resume_state(path)
    assert resumed_state["iteration"] == 4 and len(resumed_state["pairs"]) == 3
    second = run(optimize(None, None, None, 0.0, None, objective,
                          OptimizationControls(max_iterations=10, gradient_tol=1e-12, history=3),
                          resume_state=resumed_state, stationary=stationary))
    a, b = full[-1]["state"], second[-1]["state"]
    assert a["iteration"] == b["iteration"] == 10
    assert a["evaluations"] == b["evaluations"]
    for key in ("R", "K", "rho", "gradient"):
        assert np.array_equal(a[key], b[key])
    assert a["energy"] == b["energy"]


def test_noisy_energy_armijo_stagnates_wolfe_converges():
    base = OptimizationControls(max_iterations=400, gradient_tol=1e-10, history=5)
    r, k, rho, e, g, objective, stationary = quadratic_problem(noise=1e-9, power=2)
    armijo = run(optimize(r, k, rho, e, g, objective, base, robustness=RobustnessControls(), stationary=stationary))
    stop = armijo[-1]
    print(f"armijo: {stop['reason']} at |g| = {stop['state']['gradient_norm']:.2e}", file=sys.stderr)
    assert not stop["converged"] and "stagnation" in stop["reason"]
    assert stop["state"]["gradient_norm"] > 100 * base.gradient_tol
    wolfe = run(optimize(r, k, rho, e, g, objective, base, stationary=stationary,
                         robustness=RobustnessControls(line_search="wolfe", wolfe_energy_atol=1e-8)))
    stop = wolfe[-1]
    print(f"wolfe: {stop['reason']} at |g| = {stop['state']['gradient_norm']:.2e}", file=sys.stderr)
    assert stop["converged"] and stop["state"]["gradient_norm"] <= 1e-10
    assert any(ev.get("wolfe_condition") == "approximate" for ev in wolfe if ev["type"] == "accepted")


def test_plain_quadratic_converges_and_budget_stops_resumably():
    r, k, rho, e, g, objective, stationary = quadratic_problem()
    done = run(optimize(r, k, rho, e, g, objective, OptimizationControls(gradient_tol=1e-10), stationary=stationary))
    assert done[-1]["converged"] and done[-1]["state"]["gradient_norm"] <= 1e-10
    budget = run(optimize(r, k, rho, e, g, objective, OptimizationControls(gradient_tol=1e-12, max_seconds=1e-9),
                          stationary=stationary))
    assert budget[-1]["reason"].startswith("time budget max_seconds spent")
    assert budget[-1]["state"]["iteration"] == 0
