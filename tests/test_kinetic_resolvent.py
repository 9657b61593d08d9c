"""Oracle of the infinite-line kinetic backend (src/kinetic.jl), the dense-Schur angular resolvent.

Pinned the way the reference's own tests do (test/runtests.jl:261-290, test/kinetic_krylov.jl):
the coherent closed forms (kinetic1 = 0.1, kinetic2 = 0 at R = 0.3 + 0.1i); agreement with the
finite-window ODE kinetic terms when the window is wide (max_point 30: energy 1e-10, gradient
2e-8, here observed 5e-14 / 3e-13), which checks value, covector and environment against the
independently pinned ODE path; central differences of the resolvent-backend fg_phi4 along the
retraction; and failure when node doubling cannot converge.
"""
import numpy as np
import pytest

import oracle
from conftest import random_state_arrays


def prepared(d, seed, scaling):
    r, k = random_state_arrays(d, seed, scaling)
    q = oracle.canonical_q(r, k)
    return r, k, q, oracle.steady_state(q, r, tol=1e-13)


def test_coherent_closed_forms():
    R = np.array([[0.3 + 0.1j]])
    Q = oracle.canonical_q(R, np.zeros((1, 1)))
    v, cov, env, used = oracle.kinetic_resolvent(R, Q, np.ones((1, 1), complex))
    assert v == pytest.approx(0.1, abs=1e-14)
    assert used == 64


def test_resolvent_matches_wide_window_ode():
    r, k, q, rho = prepared(3, 90, 0.2)
    kw = dict(n_quad=256, max_point=30.0, atol=1e-13, rtol=1e-12)
    ode = oracle.fg_phi4(r, q, rho, 2.0, **kw)
    res = oracle.fg_phi4(r, q, rho, 2.0, kinetic_backend="resolvent", **kw)
    assert res["energy"] == pytest.approx(ode["energy"], abs=1e-10)
    np.testing.assert_allclose(res["gradient"], ode["gradient"], atol=2e-8, rtol=2e-8)
    v, _, _, _ = oracle.kinetic_resolvent(r, q, rho)
    assert v == pytest.approx(ode["term_values"][0] + ode["term_values"][1], abs=1e-10)
    # the quartic term is unchanged, the kinetic term is the resolvent value
    assert res["term_values"][2] == pytest.approx(ode["term_values"][2], rel=1e-13)


def test_resolvent_gradient_by_central_differences():
    r, k, q, rho = prepared(2, 702, 0.3)
    w = np.random.default_rng(103).standard_normal((2, 2)) + 1j * np.random.default_rng(104).standard_normal((2, 2))
    w /= np.linalg.norm(w)
    kw = dict(n_quad=128, max_point=20.0, atol=1e-13, rtol=1e-12, kinetic_backend="resolvent")
    full = oracle.fg_phi4(r, q, rho, 2.0, **kw)

    def energy(t):
        rr, kk = oracle.retraction(r, k, w, t)
        qq = oracle.canonical_q(rr, kk)
        return oracle.fg_phi4(rr, qq, oracle.steady_state(qq, rr, tol=1e-14), 2.0, **kw)["energy"]

    h = 1e-4
    fd = (-energy(2 * h) + 8 * energy(h) - 8 * energy(-h) + energy(-2 * h)) / (12 * h)
    assert oracle.metric(rho, full["gradient"], w) == pytest.approx(fd, abs=1e-8)


def test_node_doubling_failure():
    r, k, q, rho = prepared(3, 90, 0.2)
    with pytest.raises(oracle.OracleFailure):
        oracle.kinetic_resolvent(r, q, rho, minnodes=2, maxnodes=4, atol=1e-15, rtol=0.0)
