"""The Euclidean gradient of the energy in (R, K) (the backward of the autograd Function).

The oracle's gradient is pinned by central differences of its own energy along general
directions (dR arbitrary, dK Hermitian, rho0 re-solved), and by the identity that along a
horizontal direction W (dK = i/2 (W^+ R - R^+ W), the retraction of src/optimization.jl:7-13) it
reproduces metric(G, W) of the Riemannian gradient (src/observables.jl:127-130,201-212), which the
reference's own finite-difference test pins (test/runtests.jl:106-149).
"""
import numpy as np
import pytest

import oracle
from conftest import random_state_arrays

KW = dict(max_point=12.0, atol=2e-12, rtol=2e-12)


def small_state():
    base_r, base_k = random_state_arrays(3, 182, 1.0)
    return 0.12 * base_r, 0.12**2 * base_k


def energy(r, k, g, **kw):
    q = oracle.canonical_q(r, k)
    rho = oracle.steady_state(q, r, tol=1e-13)
    return oracle.fg_phi4(r, q, rho, g, boundary=False, **kw)["energy"]


def pairing(GR, GK, dr, dk):
    return float(np.vdot(GR, dr).real + np.vdot(GK, dk).real)


def directions(d, seed):
    rng = np.random.default_rng(seed)
    dr = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    x = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    dk = x + x.conj().T
    return dr / np.linalg.norm(dr), dk / np.linalg.norm(dk)


@pytest.mark.parametrize("which", ["R", "K", "both"])
def test_rk_gradient_by_central_differences(which):
    r, k = small_state()
    g = 1.2
    e, GR, GK = oracle.fg_phi4_rk(r, k, g, steady_tol=1e-13, **KW)
    np.testing.assert_allclose(GK, GK.conj().T, atol=1e-14)
    dr, dk = directions(3, {"R": 1, "K": 2, "both": 3}[which])
    if which == "R":
        dk = 0 * dk
    if which == "K":
        dr = 0 * dr
    h = 1e-5
    fd = (energy(r + h * dr, k + h * dk, g, **KW) - energy(r - h * dr, k - h * dk, g, **KW)) / (2 * h)
    assert pairing(GR, GK, dr, dk) == pytest.approx(fd, rel=2e-5, abs=1e-8)
    # the boundary (rho0 response) term is needed: without it the pairing misses the difference
    _, GR0, GK0 = oracle.fg_phi4_rk(r, k, g, boundary=False, **KW)
    assert abs(pairing(GR0, GK0, dr, dk) - fd) > 1e-4


def test_rk_gradient_reduces_to_riemannian_on_horizontal_directions():
    r, k = small_state()
    g = 1.2
    q = oracle.canonical_q(r, k)
    rho = oracle.steady_state(q, r, tol=1e-13)
    e, GR, GK = oracle.fg_phi4_rk(r, k, g, rho=rho, **KW)
    full = oracle.fg_phi4(r, q, rho, g, metric_reg=0.0, **KW)
    assert e == full["energy"]
    for seed in range(3):
        w = directions(3, 10 + seed)[0]
        dk = 0.5j * (w.conj().T @ r - r.conj().T @ w)
        assert pairing(GR, GK, w, dk) == pytest.approx(oracle.metric(rho, full["gradient"], w, 0.0), rel=1e-10)
