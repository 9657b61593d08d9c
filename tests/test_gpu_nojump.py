"""The no-jump preconditioner on defective and dark states, through the C ABI.

The stationary-state, boundary and excitation Poisson solves are preconditioned by the no-jump
inverse (src/transfer.jl:50-68).  The device applies it in the eigenbasis of Q unless Q's
eigenvectors are ill-conditioned, and then in the Schur basis (the reference's own form), with
the triangular Sylvester system solved on the device (rc_sylvester_tri).  The reference tests a
defective no-jump matrix and a dark state (test/runtests.jl:74-87); here:

- the dark state R = [0 1; 0 0], K = 0 has rho0 = diag(1, 0) (1e-10), and a given rho0 that is not a
  density fails with SolverFailure;
- canonical states whose Q is one Jordan block (D = 2, 4, 8: eigenvector condition 3e7 to 1e13,
  built from Q = -mu I + N and R^+R = -(Q + Q^+), rotated by a seeded unitary) take the Schur form, and
  their rho0 (1e-10), boundary solve (1e-9) and fg_phi4 at D = 8 (energy 1e-12 relative, gradient
  1e-10 relative to max |G|) match the oracle, whose no-jump inverse is a dense Schur solve;
- the Schur form forced on ordinary states (D = 16, 40, 64) gives the eigenbasis results to the solver
  tolerances, and the excitation Poisson solves of a D = 4 state in the Schur form match the oracle
  (1e-11).
"""
import sys

import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from oracle import excitation as X

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rg():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import rcmps_gpu
    yield rcmps_gpu
    rcmps_gpu.nojump_threshold(1e6)


def jordan_state(d, seed, c=1.0):
    """R, K with Q = -iK - R^+R/2 unitarily similar to -mu I + c N (one Jordan block)."""
    rng = np.random.default_rng(seed)
    n = np.diag(np.full(d - 1, c), 1).astype(complex)
    h = n + n.conj().T
    mu = np.linalg.norm(h, 2) / 2 + 0.1
    q = -mu * np.eye(d) + n
    k = 1j * (q - q.conj().T) / 2
    u, _ = np.linalg.qr(rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d)))
    v, _ = np.linalg.qr(rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d)))
    r = u @ sl.sqrtm(2 * mu * np.eye(d) - h)
    r, k = v @ r @ v.conj().T, v @ k @ v.conj().T
    return r, (k + k.conj().T) / 2


def hermitian(d, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    return a + a.conj().T


def test_dark_state(rg):
    note("case: dark state R=[0 1; 0 0], K=0")
    r = np.array([[0, 1], [0, 0]], dtype=complex)
    s = rg.State(r, np.zeros((2, 2)))
    assert np.abs(s.rho - np.diag([1, 0])).max() < 1e-10
    note(f"  nojump form {s.nojump_form}")
    s.close()
    with pytest.raises(rg.SolverFailure):
        rg.State(r, np.zeros((2, 2)), rho=np.diag([-1.0, 2.0]).astype(complex))


@pytest.mark.parametrize("d", [2, 4, 8])
def test_defective_q_takes_the_schur_form(rg, d):
    note(f"case: Jordan-block Q, D={d}")
    r, k = jordan_state(d, d)
    s = rg.State(r, k, steady_tol=1e-12)
    form, kappa = s.nojump_form
    note(f"  nojump form {form}, kappa {kappa:.3e}, rho0 matvecs {s.steady_stats['krylov_matvecs']}")
    assert form == "schur" and kappa > 1e6
    np.testing.assert_allclose(s.Q, oracle.canonical_q(s.R, s.K), atol=1e-14)
    ref = oracle.steady_state(s.Q, s.R, method="dense")
    assert np.abs(s.rho - ref).max() < 1e-10
    a = hermitian(d, 12)
    h = rg.boundary_environment(s, a, tol=1e-12)
    href = oracle.boundary_environment(s.Q, s.R, s.rho, a, method="dense")
    assert np.abs(h - href).max() < 1e-9
    assert abs(np.trace(s.rho @ h)) < 1e-9
    if d == 8:
        e, g = rg.fg_phi4(s, 2.0)
        o = oracle.fg_phi4(s.R, s.Q, s.rho, 2.0)
        note(f"  E = {e!r}, oracle {o['energy']!r}")
        assert abs(e - o["energy"]) <= 1e-12 * abs(o["energy"])
        assert np.abs(g - o["gradient"]).max() <= 1e-10 * np.abs(o["gradient"]).max()
    s.close()


@pytest.mark.parametrize("d", [16, 40, 64])
def test_forced_schur_form_equals_eigenbasis(rg, d):
    note(f"case: Schur form forced, D={d}")
    r, k = rg.random_state_arrays(d, scaling=0.2, seed=42)
    a = hermitian(d, 3)
    out = {}
    for kappa in (1e6, 0.0):
        rg.nojump_threshold(kappa)
        s = rg.State(r, k, steady_tol=1e-12)
        form, est = s.nojump_form
        h = rg.boundary_environment(s, a, tol=1e-12)
        e, g = rg.fg_phi4(s, 2.0) if d <= 40 else (None, None)
        out[form] = (s.rho, h, e, g)
        note(f"  {form}: kappa {est:.3e}, rho0 matvecs {s.steady_stats['krylov_matvecs']}")
        s.close()
    rg.nojump_threshold(1e6)
    (r1, h1, e1, g1), (r0, h0, e0, g0) = out["eigen"], out["schur"]
    assert np.abs(r1 - r0).max() < 1e-11
    assert np.abs(h1 - h0).max() < 1e-9 * max(1.0, np.abs(h1).max())
    if e1 is not None:
        assert abs(e1 - e0) <= 1e-11 * abs(e1)
        assert np.abs(g1 - g0).max() <= 1e-9 * np.abs(g1).max()


def test_excitation_poisson_in_the_schur_form(rg):
    note("case: excitation Poisson solves, Schur form, D=4 p=0.7")
    from rcmps_gpu.excitation import ExcitationContext, action_dual_poisson, excitation_poisson
    rng = np.random.default_rng(3)
    x = 0.5 * (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))
    k = (x + x.conj().T) / 2
    r = 0.5 * (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))
    rg.nojump_threshold(0.0)
    try:
        s = rg.State(r, k, steady_tol=1e-13)
    finally:
        rg.nojump_threshold(1e6)
    assert s.nojump_form[0] == "schur"
    ctx = ExcitationContext(s, 0.7, dense_transfer=False)   # the GMRES route, preconditioned in the Schur form
    ip = s.integrator_param
    octx = X.context(X.XState(s.R, s.Q, s.rho, n_quad=ip.n_quad, max_point=ip.max_point, atol=ip.atol,
                              rtol=ip.rtol), 0.7)
    b = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    for shift in (-1, 0, 1):
        assert np.abs(excitation_poisson(ctx, b, shift) - X.excitation_poisson(octx, b, shift)).max() < 1e-11
    for shifted in (False, True):
        assert np.abs(action_dual_poisson(ctx, b, shifted) - X.action_dual_poisson(octx, b, shifted)).max() < 1e-11
    ctx.close()
    s.close()
