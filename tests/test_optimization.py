"""The OptimKit 0.5 algorithms behind optimize_state (rcmps_gpu.optimization.optimkit_optimize) on
the CPU, on quadratics in R^n with a diagonal preconditioner.

- The L-BFGS inverse Hessian of two normalised pairs, applied by the two-loop recursion with the
  secant-scaled preconditioner, equals the matrix BFGS recursion H <- (I - r s y^T) H (I - r y s^T)
  + r s s^T started from H0 = <s,y>/<y,P y> P of the newest pair.
- The first L-BFGS step is -0.01 P g / |P g| tried at alpha = 1 with acceptfirst; GD and CG start at
  alpha = 1/|P g| and double the accepted step; GD and CG searches do not accept first trials.
- With exact line searches on an n-dimensional quadratic, CG with the Hager-Zhang beta and L-BFGS
  with memory >= n both converge in n steps (the conjugate-direction property: a dropped or wrongly
  signed term of beta, or a wrong scaling, loses it), and GD does not.
"""
import math

import numpy as np
import pytest

from rcmps_gpu.optimization import hager_zhang_beta, lbfgs_inverse_hessian, optimkit_optimize


def quadratic(n=6, seed=3):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    a = q @ np.diag(np.linspace(1.0, 12.0, n)) @ q.T
    b = rng.standard_normal(n)
    p_diag = rng.uniform(0.5, 2.0, n)

    def fg(x):
        return 0.5 * x @ a @ x - b @ x, a @ x - b

    def exact_ls(fg_, x, p, fg0, initialguess=1.0, *, retract, inner, acceptfirst=False):
        g0 = fg0[1]
        alpha = -(g0 @ p) / (p @ a @ p)
        y, v = retract(x, p, alpha)
        f, g = fg_(y)
        return y, f, g, v, alpha, 1

    ops = dict(inner=lambda x, u, w: float(u @ w), precondition=lambda x, g: p_diag * g,
               retract=lambda x, p, alpha: (x + alpha * p, p), finalize=lambda x, f, g: x)
    return a, b, p_diag, fg, exact_ls, ops


def test_two_pair_inverse_hessian_equals_the_bfgs_recursion():
    rng = np.random.default_rng(7)
    n = 5
    p_diag = rng.uniform(0.5, 2.0, n)
    m = rng.standard_normal((n, n))
    a = m @ m.T + n * np.eye(n)
    raw = []
    for _ in range(2):
        s = rng.standard_normal(n)
        raw.append((s, a @ s))
    pairs = [(s / np.linalg.norm(s), y / np.linalg.norm(s), (s @ s) / (s @ y)) for s, y in raw]
    g = rng.standard_normal(n)
    got = lbfgs_inverse_hessian(g, pairs, lambda u, w: float(u @ w), lambda v: p_diag * v)
    s_m, y_m = raw[-1]
    h = (s_m @ y_m) / (y_m @ (p_diag * y_m)) * np.diag(p_diag)
    for s, y in raw:
        r = 1.0 / (s @ y)
        left = np.eye(n) - r * np.outer(s, y)
        h = left @ h @ left.T + r * np.outer(s, s)
    np.testing.assert_allclose(got, h @ g, rtol=1e-12, atol=1e-12)


def test_first_steps_and_initial_guesses():
    a, b, p_diag, fg, exact_ls, ops = quadratic()
    x0 = np.zeros(len(b))
    for method in ("LBFGS", "GD", "CG"):
        calls = []

        def recording_ls(fg_, x, p, fg0, initialguess=1.0, **kw):
            calls.append((p.copy(), initialguess, kw["acceptfirst"]))
            return exact_ls(fg_, x, p, fg0, initialguess, **kw)

        optimkit_optimize(fg, x0, method, ls=recording_ls, n_iter=3, gradtol=1e-14, **ops)
        g0 = -b
        pg = p_diag * g0
        p0, guess0, first0 = calls[0]
        if method == "LBFGS":
            np.testing.assert_allclose(p0, -0.01 * pg / math.sqrt(pg @ pg), rtol=1e-14)
            assert [c[1] for c in calls] == [1.0, 1.0, 1.0] and all(c[2] for c in calls)
        else:
            np.testing.assert_allclose(p0, -pg, rtol=1e-14)
            assert guess0 == pytest.approx(1.0 / math.sqrt(pg @ pg), rel=1e-14)
            assert not any(c[2] for c in calls)
            # the next initial guess doubles the accepted step
            x1 = x0 + (-(g0 @ p0) / (p0 @ a @ p0)) * p0
            assert calls[1][1] == pytest.approx(2 * (-(g0 @ p0) / (p0 @ a @ p0)), rel=1e-14)
            assert np.isfinite(x1).all()


@pytest.mark.parametrize("method", ["CG", "LBFGS"])
def test_conjugate_directions_converge_in_n_steps(method):
    a, b, p_diag, fg, exact_ls, ops = quadratic()
    n = len(b)
    x, f, g, evaluations, history = optimkit_optimize(fg, np.zeros(n), method, ls=exact_ls, n_iter=n,
                                                      gradtol=1e-300, mlbfgs=n, **ops)
    assert len(history) == n + 1
    np.testing.assert_allclose(x, np.linalg.solve(a, b), rtol=1e-9, atol=1e-9)
    assert history[-1][1] < 1e-9 * history[0][1]


def test_gradient_descent_is_not_conjugate():
    a, b, p_diag, fg, exact_ls, ops = quadratic()
    n = len(b)
    _, _, _, _, history = optimkit_optimize(fg, np.zeros(n), "GD", ls=exact_ls, n_iter=n, gradtol=1e-300, **ops)
    assert history[-1][1] > 1e-6 * history[0][1]


def test_hager_zhang_beta_lower_bound_and_unknown_method():
    inner = lambda u, w: float(u @ w)  # noqa: E731
    # a step where the unbounded coefficient (-1.59) lies below the bound 0.4 <d,g_prev>/<d,d> (-0.095)
    d = np.array([0.5, 1.6])
    g_prev = np.array([-2.3, 0.3])
    g = np.array([-1.1, 0.6])
    beta = hager_zhang_beta(g, g_prev, g, g_prev, d, inner)
    assert beta == pytest.approx(0.4 * inner(d, g_prev) / inner(d, d))
    with pytest.raises(ValueError):
        optimkit_optimize(lambda x: (0.0, x), np.ones(2), "BFGS", inner=None, precondition=None, retract=None,
                          finalize=None, ls=None, n_iter=1, gradtol=1.0)
