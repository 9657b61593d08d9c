"""optimize_state on the GPU with the default SafeguardedLineSearch and with Armijo backtracking.

Mirrors the reference's optimizer tests (test/runtests.jl:166-172, 195-197, 201-219): two
iterations of GD and L-BFGS lower the energy; the default Wolfe search lowers it too; failed
trials (SolverFailure raised by the objective) are rejected by both line searches, and the
safeguarded one keeps |alpha p| within its relative step bound.  Parity with the oracle: three
L-BFGS iterations driven by the GPU fg_phi4 and by the oracle's fg_phi4 on the same states take
the same line-search steps and reach energies within 1e-10 of each other.
"""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rg():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import rcmps_gpu
    return rcmps_gpu


def small_state(rg):
    s = rg.RandomState(3, seed=16, scaling=0.3)
    s.integrator_param.atol = 1e-12
    s.integrator_param.rtol = 1e-12
    s.integrator_param.n_quad = 64
    s.model_param.name = "phi4"
    s.model_param.coupling = 1.0
    return s


@pytest.mark.parametrize("method", ["GD", "LBFGS"])
def test_two_iterations_with_backtracking_lower_the_energy(rg, method):
    note(f"case: optimize_state {method} D=3 n_iter=2 BackwardLineSearch")
    s = small_state(rg)
    initial, _ = rg.fg_phi4(s, 1.0)
    result = rg.optimize_state(s, method=method, n_iter=2, ls=rg.BackwardLineSearch(), diagnostics=True)
    assert result["energy"] < initial
    assert np.isfinite(result["gradient_norm"])
    with pytest.raises(ValueError):
        rg.optimize_state(s, method="bogus")


def test_default_wolfe_search_lowers_the_energy(rg):
    note("case: optimize_state LBFGS D=3 n_iter=2 SafeguardedLineSearch")
    s = small_state(rg)
    initial, _ = rg.fg_phi4(s, 1.0)
    ls = rg.SafeguardedLineSearch()
    result = rg.optimize_state(s, n_iter=2, ls=ls, diagnostics=True)
    assert result["energy"] < initial
    assert ls.last_used == "wolfe"
    assert len(result["history"]) == 3


def test_line_searches_reject_failed_trials(rg):
    note("case: line searches with failing trials D=1")
    from rcmps_gpu import SolverFailure, metric, retraction
    s = rg.State(np.array([[0.3 + 0j]]), np.zeros((1, 1)))
    attempts = [0]

    def objective(x):
        attempts[0] += 1
        if abs(x.R[0, 0]) > 0.5:
            raise SolverFailure("synthetic invalid trial")
        return abs(x.R[0, 0]) ** 2, x.R / (1 + x.manifold_param.metric_reg)

    f, g = objective(s)
    direction = -4 * g
    y, fy, gy, v, alpha, nfg = rg.BackwardLineSearch()(objective, s, direction, (f, g), retract=retraction,
                                                       inner=metric)
    assert fy < f
    assert attempts[0] > 2
    assert np.abs(y.Q + y.Q.conj().T + y.R.conj().T @ y.R).max() < 1e-14
    ls = rg.SafeguardedLineSearch(max_relative_step=0.05)
    y, fy, gy, v, alpha, nfg = ls(objective, s, direction, (f, g), retract=retraction, inner=metric)
    assert fy < f
    assert np.linalg.norm(alpha * direction) <= 0.05


def test_lbfgs_steps_match_the_oracle(rg, monkeypatch):
    note("case: optimize_state LBFGS D=4 n_iter=3, GPU fg_phi4 against the oracle's")
    import rcmps_gpu.optimization as opt

    def make():
        s = rg.RandomState(4, seed=7, scaling=0.3)
        s.model_param.name = "phi4"
        s.model_param.coupling = 2.0
        s.integrator_param.n_quad = 64
        return s

    runs = {}
    for use_oracle in (False, True):
        calls = []
        if use_oracle:
            def energy_function(s):
                ip = s.integrator_param

                def fg(x):
                    d = oracle.fg_phi4(x.R, x.Q, x.rho, 2.0, n_quad=ip.n_quad, max_point=ip.max_point,
                                       atol=ip.atol, rtol=ip.rtol, metric_reg=x.manifold_param.metric_reg)
                    calls.append(d["energy"])
                    return d["energy"], d["gradient"]
                return fg
        else:
            base = opt.energy_function

            def energy_function(s):
                inner = base(s)

                def fg(x):
                    e, G = inner(x)
                    calls.append(e)
                    return e, G
                return fg
        monkeypatch.setattr(opt, "energy_function", energy_function)
        res = rg.optimize_state(make(), n_iter=3, diagnostics=True)
        runs[use_oracle] = (res, calls)
        note(f"  {'oracle' if use_oracle else 'gpu'}: evaluations={res['evaluations']} energy={res['energy']:.15f}")
    (a, ca), (b, cb) = runs[False], runs[True]
    assert a["evaluations"] == b["evaluations"] and len(ca) == len(cb)
    assert np.abs(np.array(ca) - np.array(cb)).max() < 1e-10
    assert a["energy"] < ca[0]


def test_controller_resume_on_the_gpu(rg, tmp_path):
    """The resumable controller with fg_phi4 and the stationary state on the GPU: energies fall at
    every accepted step, and four iterations equal two, a saved snapshot, and two more, bit for bit."""
    note("case: controller D=8 g=2 seed=5, 4 iterations vs 2 + resume 2")
    from rcmps_gpu.controller import OptimizationControls, RobustnessControls, optimize
    r, k = rg.random_state_arrays(8, scaling=0.2, seed=5)
    s = rg.State(r, k, steady_tol=1e-13)
    ip = rg.IntegratorParameters(n_quad=64, max_point=20.0, atol=1e-11, rtol=1e-10)
    s.integrator_param = ip
    e0, g0 = rg.fg_phi4(s, 2.0)
    objective, stationary = rg.gpu_objective(2.0, integrator=ip), rg.gpu_stationary()
    robust = RobustnessControls(line_search="wolfe", wolfe_energy_atol=1e-10)

    def events(n, resume=None):
        out = list(optimize(s.R, s.K, s.rho, e0, g0, objective, OptimizationControls(max_iterations=n, history=5),
                            resume_state=resume, robustness=robust, stationary=stationary))
        assert out[-1]["type"] == "stopped"
        return out

    full = events(4)
    energies = [e0] + [ev["energy"] for ev in full if ev["type"] == "accepted"]
    note(f"  energies {energies}")
    assert len(energies) == 5 and energies[-1] < energies[0]
    half = events(2)
    path = rg.save_resume_state(tmp_path / "state.npz", half[-1]["state"])
    rest = events(4, rg.load_resume_state(path))
    a, b = full[-1]["state"], rest[-1]["state"]
    for key in ("R", "K", "rho", "gradient"):
        assert np.array_equal(a[key], b[key])
    assert a["energy"] == b["energy"] and a["evaluations"] == b["evaluations"]
    s.close()
