"""The Euclidean (R, K) gradient on the GPU and the torch.autograd.Function built on it.

Parity: energy within 1e-12 relative and both gradients within 1e-10 relative (to their max
entry) of the oracle's fg_phi4_rk on the same state and rho0; the energy-only call (forward
trajectories only) returns the same energy.  The D = 64, g = 1 differentiation workload: central
differences of the energy with step 1e-6 along 16 seeded complex directions (dR arbitrary, dK
Hermitian) agree with the autograd gradient within 1e-5 relative; then 50 L-BFGS iterations in
(R, K) through the autograd Function descend at every step and lower the gradient norm.
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


def directions(d, seed):
    rng = np.random.default_rng(seed)
    dr = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    x = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    dk = x + x.conj().T
    return dr / np.linalg.norm(dr), dk / np.linalg.norm(dk)


@pytest.mark.parametrize("d,seed,scaling,n_quad", [(4, 7, 0.3, 32), (8, 42, 0.2, 64), (16, 11, 0.2, 64)])
def test_rk_gradient_matches_oracle(rg, d, seed, scaling, n_quad):
    note(f"case: fg_phi4_rk D={d} seed={seed} n_quad={n_quad}")
    r, k = rg.random_state_arrays(d, scaling=scaling, seed=seed)
    s = rg.State(r, k)
    s.integrator_param.n_quad = n_quad
    e, GR, GK, st = rg.fg_phi4_rk(s, 2.0)
    GR, GK = GR.cpu().numpy(), GK.cpu().numpy()
    eo, GRo, GKo = oracle.fg_phi4_rk(s.R, s.K, 2.0, rho=s.rho, n_quad=n_quad)
    assert e == pytest.approx(eo, rel=1e-12)
    assert np.abs(GR - GRo).max() <= 1e-10 * np.abs(GRo).max()
    assert np.abs(GK - GKo).max() <= 1e-10 * np.abs(GKo).max()
    e_only, none_r, none_k, st1 = rg.fg_phi4_rk(s, 2.0, gradient=False)
    assert none_r is None and none_k is None
    assert e_only == pytest.approx(e, rel=1e-14)
    assert all(st1["accepted"][i] == 0 for i in (1, 3, 5))
    s.close()


def test_autograd_function_backward(rg):
    import torch
    note("case: Phi4Energy backward D=8")
    r, k = rg.random_state_arrays(8, scaling=0.2, seed=5)
    R = torch.tensor(r, requires_grad=True)
    K = torch.tensor(k, requires_grad=True)
    E = rg.phi4_energy(R, K, 2.0)
    (3.0 * E).backward()
    eo, GRo, GKo = oracle.fg_phi4_rk(r, k, 2.0, steady_tol=1e-12)
    assert float(E) == pytest.approx(eo, rel=1e-10)
    np.testing.assert_allclose(R.grad.numpy(), 3.0 * GRo, atol=1e-8 * np.abs(GRo).max())
    np.testing.assert_allclose(K.grad.numpy(), 3.0 * GKo, atol=1e-8 * np.abs(GKo).max())


def test_d64_central_differences_16_directions(rg):
    import torch
    from rcmps_gpu import IntegratorParameters
    note("case: D=64 g=1 central differences, 16 directions, step 1e-6")
    ip = IntegratorParameters(atol=1e-13, rtol=1e-12)
    r, k = rg.random_state_arrays(64, scaling=0.2, seed=42)
    dev = torch.device("cuda", 0)
    R = torch.tensor(r, device=dev, requires_grad=True)
    K = torch.tensor(k, device=dev, requires_grad=True)
    E = rg.phi4_energy(R, K, 1.0, integrator=ip, steady_tol=1e-13)
    E.backward()
    GR, GK = R.grad.cpu().numpy(), K.grad.cpu().numpy()
    h = 1e-6
    worst = 0.0
    with torch.no_grad():
        for i in range(16):
            dr, dk = directions(64, 100 + i)
            ep = rg.phi4_energy(torch.tensor(r + h * dr, device=dev), torch.tensor(k + h * dk, device=dev), 1.0,
                                integrator=ip, steady_tol=1e-13)
            em = rg.phi4_energy(torch.tensor(r - h * dr, device=dev), torch.tensor(k - h * dk, device=dev), 1.0,
                                integrator=ip, steady_tol=1e-13)
            fd = (float(ep) - float(em)) / (2 * h)
            an = float(np.vdot(GR, dr).real + np.vdot(GK, dk).real)
            err = abs(fd - an) / max(abs(an), 1e-3)
            note(f"  direction {i}: fd={fd:.12e} adjoint={an:.12e} rel={err:.2e}")
            worst = max(worst, err)
    assert worst < 1e-5


def test_d64_lbfgs_50_iterations(rg):
    note("case: D=64 g=1 L-BFGS in (R, K), 50 iterations")
    from rcmps_gpu import lbfgs_rk, phi4_energy
    import torch
    r, k = rg.random_state_arrays(64, scaling=0.2, seed=42)
    R, K, energies, evaluations = lbfgs_rk(r, k, 1.0, n_iter=50)
    note(f"  iterations={len(energies) - 1} evaluations={evaluations} first={energies[0]:.12f} "
         f"last={energies[-1]:.12f}")
    assert len(energies) == 51
    assert all(b <= a for a, b in zip(energies, energies[1:]))   # Armijo: every iteration descends
    assert energies[-1] < energies[0] - 1e-3
    dev = torch.device("cuda", 0)
    grads = []
    for rr, kk in ((r, k), (R, K)):
        Rt = torch.tensor(rr, device=dev, requires_grad=True)
        Kt = torch.tensor(kk, device=dev, requires_grad=True)
        phi4_energy(Rt, Kt, 1.0).backward()
        grads.append(float(torch.sqrt(Rt.grad.abs().square().sum() + Kt.grad.abs().square().sum())))
    note(f"  gradient norm {grads[0]:.3e} -> {grads[1]:.3e}")
    assert grads[1] < grads[0]
