"""Pins the CPU oracle against known answers, closed forms and independent mathematics.

Each check is chosen so that a plausible slip in the oracle (a dropped term, a wrong sign, a
transposed operand, a wrong tableau entry) fails it.  Reference tests restated here:
test/runtests.jl:22-50 (canonical identities, vec/transfer, adjoint duality), :52-88
(stationary state, boundary Poisson, defective no-jump matrix, dark state), :90-104 (closed
forms), :106-149 (finite differences of the gradient), :175-198 (n_quad refinement),
:220-259 (threaded == serial).
"""
import math
from pathlib import Path

import mpmath
import numpy as np
import pytest

import oracle
from conftest import FIXTURES, random_state_arrays


def canonical(d, seed, scaling):
    r, k = random_state_arrays(d, seed, scaling)
    q = oracle.canonical_q(r, k)
    return r, k, q


def colvec(x):
    return x.reshape(-1, order="F")


def transfer_matrix(q, r):
    """src/transfer.jl:21-25 (Julia column-major vec)."""
    i = np.eye(len(q))
    return np.kron(i, q) + np.kron(q.conj(), i) + np.kron(r.conj(), r)


def test_source_against_mpmath():
    assert oracle.source(0.0) == pytest.approx(1 / math.sqrt(math.pi), abs=1e-16)
    assert oracle.source(1e-10) == pytest.approx(1 / math.sqrt(math.pi), rel=1e-8)
    c = 2**2.25 * math.sqrt(math.pi) * float(mpmath.gamma(1.25))
    for t in np.linspace(0.003, 6.3, 41):
        x = t * t
        ref = float(2 * t * mpmath.besselk(0.25, x) / (c * mpmath.mpf(x) ** 0.25))
        assert oracle.source(t) == pytest.approx(ref, rel=4e-15, abs=1e-300)
        assert oracle.source(-t) == oracle.source(t)


def test_source_integral_over_line():
    """int J over R = 2^-1/2; on the window the tail is ~e^-W (src/observables.jl:1-5)."""
    p, w = oracle.quadrature(256, 30.0)
    s = np.array([oracle.source(t) for t in p])
    # int_{-a}^{a} 2|t| J(t^2) dt = int_{-W}^{W} J(x) dx
    assert np.sum(w * s) == pytest.approx(2**-0.5, abs=1e-12)


def test_quadrature_matches_gauss_legendre():
    for n, mp in ((8, 20.0), (128, 20.0), (192, 30.0)):
        p, w = oracle.quadrature(n, mp)
        x, ww = np.polynomial.legendre.leggauss(n)
        a = math.sqrt(mp)
        np.testing.assert_allclose(p[n:], a / 2 * (x + 1), rtol=0, atol=2e-15)
        # numpy's end weights carry ~3e-11 relative error; 40-digit values pin the extremes
        np.testing.assert_allclose(w[n:], a / 2 * ww, rtol=1e-10, atol=0)
        mpmath.mp.dps = 40
        for i in (0, 1, n // 2):
            z = mpmath.findroot(lambda s: mpmath.legendre(n, s), x[i])
            dp = mpmath.diff(lambda s: mpmath.legendre(n, s), z)
            ref = float(a / 2 * 2 / ((1 - z**2) * dp**2))
            assert w[n + i] == pytest.approx(ref, rel=2e-12)
        mpmath.mp.dps = 15
        np.testing.assert_array_equal(p[:n], -p[n:][::-1])
        np.testing.assert_array_equal(w[:n], w[n:][::-1])
        assert np.sum(w) == pytest.approx(2 * a, rel=1e-14)


def test_lindblad_transfer_and_duality():
    r, k, q = canonical(3, 61, 1.0)
    # left canonical: Q + Q^dagger + R^dagger R = 0  (test/runtests.jl:23)
    np.testing.assert_allclose(q + q.conj().T + r.conj().T @ r, 0, atol=1e-13)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    y = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    lx = oracle.lindblad(q, r, x)
    np.testing.assert_allclose(colvec(lx), transfer_matrix(q, r) @ colvec(x), atol=1e-13)
    # <y, L(x)> = <L*(y), x>
    assert np.vdot(y, lx) == pytest.approx(np.vdot(oracle.lindblad(q, r, y, adjoint=True), x), abs=1e-12)
    assert abs(np.trace(lx)) < 1e-12


def test_schur_factorisation():
    rng = np.random.default_rng(4)
    for d in (1, 2, 5, 17, 40):
        a = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
        t, z = oracle.schur(a)
        np.testing.assert_allclose(np.tril(t, -1), 0, atol=0)
        np.testing.assert_allclose(z @ t @ z.conj().T, a, atol=1e-12 * max(1, np.abs(a).max()) * d)
        np.testing.assert_allclose(z.conj().T @ z, np.eye(d), atol=1e-13 * d)
        np.testing.assert_allclose(np.sort_complex(np.diag(t)), np.sort_complex(np.linalg.eigvals(a)),
                                   atol=1e-10 * d)


def test_nojump_inverse_defective():
    """test/runtests.jl:81-87: a defective no-jump matrix."""
    q = np.array([[-1, 1], [0, -1]], dtype=complex)
    r = np.eye(2, dtype=complex)
    y = np.array([[2, 1j], [-1j, 1]])
    x = oracle.nojump(q, r, y)
    np.testing.assert_allclose(-(q @ x + x @ q.conj().T), y, atol=1e-13)
    x = oracle.nojump(q, r, y, adjoint=True)
    np.testing.assert_allclose(-(q.conj().T @ x + x @ q), y, atol=1e-13)


def test_dark_state():
    """test/runtests.jl:75-77: R=[0 1;0 0], K=0 has rho0 = diag(1,0)."""
    r = np.array([[0, 1], [0, 0]], dtype=complex)
    q = oracle.canonical_q(r, np.zeros((2, 2)))
    rho = oracle.steady_state(q, r)
    np.testing.assert_allclose(rho, np.diag([1, 0]), atol=1e-10)


@pytest.mark.parametrize("d", [2, 5, 12])
def test_stationary_state_and_boundary(d):
    """test/runtests.jl:53-73 with the textbook dense nullspace solve as the pin."""
    r, k, q = canonical(d, d, 1.0)
    exact = oracle.steady_state(q, r, method="dense")
    it = oracle.steady_state(q, r, method="cptp")
    # independent dense nullspace of the column-major transfer matrix
    t = transfer_matrix(q, r)
    v = colvec(np.eye(d))
    ref = np.linalg.solve(t + np.outer(v, v.conj()) / d, v / d).reshape(d, d, order="F")
    np.testing.assert_allclose(exact, ref, atol=1e-11)
    assert np.linalg.norm(it - exact) < 1e-9
    assert np.linalg.norm(oracle.lindblad(q, r, it)) < 1e-8
    assert np.trace(it) == pytest.approx(1)
    assert np.linalg.eigvalsh(it).min() >= -1e-12
    a = np.random.default_rng(12).standard_normal((d, d)) * (1 + 1j)
    a = a + a.conj().T
    h = oracle.boundary_environment(q, r, exact, a)
    hd = oracle.boundary_environment(q, r, exact, a, method="dense")
    np.testing.assert_allclose(h, hd, atol=1e-8, rtol=1e-8)
    assert abs(np.trace(exact @ h)) < 1e-9
    np.testing.assert_allclose(-oracle.lindblad(q, r, h, adjoint=True), a - np.trace(exact @ a) * np.eye(d),
                               atol=1e-8)
    # dense boundary against the textbook column-major system (src/transfer.jl:127-129)
    b = a - np.trace(exact @ a) * np.eye(d)
    mat = -t.conj().T + np.outer(colvec(np.eye(d)), colvec(exact).conj())
    href = np.linalg.solve(mat, colvec(b)).reshape(d, d, order="F")
    np.testing.assert_allclose(hd, href, atol=1e-10)


def test_reduced_channels_are_consistent():
    """The reduced coordinates rely on ch3 = ch2^dagger (kinetic) and hermiticity (quartic)."""
    r, k, q = canonical(4, 3, 0.5)
    rho = oracle.steady_state(q, r)
    rng = np.random.default_rng(5)
    f2 = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    # kinetic forward: channel 3 = channel 2^dagger, both fed by rho
    for term in (0, 1):
        u = np.zeros((4, 4, 4), complex)
        u[0], u[1], u[2] = rho, f2, f2.conj().T
        du = oracle.rhs_full(q, r, term, 0, 0.37, u)
        np.testing.assert_allclose(du[2], du[1].conj().T, atol=1e-13)
        # d/dt tr(F4) only has insertion terms: tr L(X) = 0 for a canonical state
        u[3] = rng.standard_normal((4, 4))
        du2 = oracle.rhs_full(q, r, term, 0, 0.37, u)
        assert np.trace(du2[3]) == pytest.approx(np.trace(du[3]), abs=1e-12)
        # adjoint: channel 3 = channel 2^dagger, channel 4 = I constant, channel 1 hermitian
        a1 = rng.standard_normal((4, 4)) * (1 + 1j)
        a1 = a1 + a1.conj().T
        ua = np.zeros((4, 4, 4), complex)
        ua[0], ua[1], ua[2], ua[3] = a1, f2, f2.conj().T, np.eye(4)
        dua = oracle.rhs_full(q, r, term, 1, -0.8, ua)
        np.testing.assert_allclose(dua[2], dua[1].conj().T, atol=1e-13)
        np.testing.assert_allclose(dua[0], dua[0].conj().T, atol=1e-13)
        np.testing.assert_allclose(dua[3], 0, atol=1e-13)
    # quartic: hermitian channels stay hermitian
    u = np.zeros((5, 4, 4), complex)
    u[0] = rho
    for c in (1, 2, 3):
        x = rng.standard_normal((4, 4)) * (1 + 1j)
        u[c] = x + x.conj().T
    du = oracle.rhs_full(q, r, 2, 0, 1.1, u)
    for c in range(5):
        np.testing.assert_allclose(du[c], du[c].conj().T, atol=1e-13)


def test_coherent_state_closed_forms():
    """test/runtests.jl:90-104: kinetic1 = 0.1, kinetic2 = 0, <:phi^4:> = (sqrt2 * 0.3)^4."""
    r = np.array([[0.3 + 0.1j]])
    q = oracle.canonical_q(r, np.zeros((1, 1)))
    out = oracle.fg_phi4(r, q, np.eye(1, dtype=complex), 1.0, max_point=30.0, atol=1e-12, rtol=1e-12)
    k1, k2, p4 = out["term_values"]
    assert k1 == pytest.approx(0.1, abs=1e-12)
    assert k2 == pytest.approx(0.0, abs=1e-12)
    assert p4 == pytest.approx((math.sqrt(2) * 0.3) ** 4, abs=1e-12)



def read_fixture():
    d = FIXTURES / "phi4_g2_D16_z2_portable"
    def ld(name):
        return np.fromfile(d / f"{name}.bin", dtype="<c16").reshape((16, 16), order="F")
    return ld("R"), ld("K"), ld("rho"), ld("gradient")


def test_golden_fixture_energy():
    """test/fixtures/phi4_g2_D16_z2_portable: energy -0.15648780475169466 at n_quad 192,
    window 20, 1e-13/1e-12, g = 2.  The oracle reproduces it to its own integration error."""
    R, K, rho, G = read_fixture()
    q = oracle.canonical_q(R, K)
    out = oracle.fg_phi4(R, q, rho, 2.0, n_quad=192, max_point=20.0, atol=1e-13, rtol=1e-12)
    e = out["energy"]
    assert abs(e - (-0.15648780475169466)) < 2e-11
    # gradient is ~3e-6 in norm: pinned absolutely, and through its metric norm
    assert np.abs(out["gradient"] - G).max() < 2e-8
    gn = math.sqrt(oracle.metric(rho, out["gradient"], out["gradient"]))
    assert gn == pytest.approx(2.7317828676883325e-06, rel=1e-3)


def test_golden_fixture_energy_tight():
    """Tightening the integrator converges onto the fixture energy (relative 1e-12)."""
    R, K, rho, G = read_fixture()
    q = oracle.canonical_q(R, K)
    out = oracle.fg_phi4(R, q, rho, 2.0, n_quad=192, max_point=20.0, atol=1e-15, rtol=1e-14)
    assert out["energy"] == pytest.approx(-0.15648780475169466, rel=1e-12)


def directional_fd(r, k, rho_fn, g, w, h=1e-5, **kw):
    def energy(rr, kk):
        q = oracle.canonical_q(rr, kk)
        rho = rho_fn(q, rr)
        return oracle.fg_phi4(rr, q, rho, g, boundary=False, **kw)["energy"]
    rp, kp = oracle.retraction(r, k, w, h)
    rm, km = oracle.retraction(r, k, w, -h)
    return (energy(rp, kp) - energy(rm, km)) / (2 * h)


def test_gradient_by_central_differences():
    """test/runtests.jl:106-149: metric(G, W) equals the central difference of E along the
    retraction; the bulk-only gradient (no boundary term) does not."""
    base_r, base_k = random_state_arrays(3, 182, 1.0)
    r, k = 0.12 * base_r, 0.12**2 * base_k
    q = oracle.canonical_q(r, k)
    rho = oracle.steady_state(q, r, tol=1e-13)
    w = np.random.default_rng(7).standard_normal((3, 3)) + 1j * np.random.default_rng(8).standard_normal((3, 3))
    w /= np.linalg.norm(w)
    kw = dict(max_point=12.0, atol=2e-12, rtol=2e-12)
    full = oracle.fg_phi4(r, q, rho, 1.2, **kw)
    bulk = oracle.fg_phi4(r, q, rho, 1.2, boundary=False, **kw)
    fd = directional_fd(r, k, lambda qq, rr: oracle.steady_state(qq, rr, tol=1e-13), 1.2, w, **kw)
    assert oracle.metric(rho, full["gradient"], w) == pytest.approx(fd, rel=2e-5, abs=1e-7)
    assert abs(oracle.metric(rho, bulk["gradient"], w) - fd) > 1e-4
    dense = oracle.fg_phi4(r, q, rho, 1.2, boundary_method="dense", **kw)
    np.testing.assert_allclose(dense["gradient"], full["gradient"], atol=1e-9)


def test_quadrature_refinement_and_threads():
    """test/runtests.jl:175-198 and :220-259."""
    r, k, q = canonical(3, 16, 0.3)
    rho = oracle.steady_state(q, r)
    kw = dict(atol=1e-12, rtol=1e-12)
    a = oracle.fg_phi4(r, q, rho, 1.0, n_quad=64, **kw)
    b = oracle.fg_phi4(r, q, rho, 1.0, n_quad=128, **kw)
    assert a["energy"] == pytest.approx(b["energy"], abs=1e-10)
    np.testing.assert_allclose(a["gradient"], b["gradient"], atol=2e-8, rtol=1e-7)
    c = oracle.fg_phi4(r, q, rho, 1.0, n_quad=128, nthreads=1, **kw)
    assert c["energy"] == b["energy"]
    np.testing.assert_array_equal(c["gradient"], b["gradient"])


def test_max_steps_failure():
    """maxiters = 1 turns into a solver failure (test/runtests.jl:147-148)."""
    r, k, q = canonical(3, 24, 0.2)
    rho = oracle.steady_state(q, r)
    with pytest.raises(oracle.OracleFailure):
        oracle.fg_phi4(r, q, rho, 2.0, max_steps=1)
