"""The phi^4 energy as a differentiable function of (R, K) (torch.autograd.Function).

E(R, K) is the fg_phi4 energy (src/observables.jl:226-257) of the left-canonical state with
Q = -iK - R^+R/2 and its stationary density rho0 solved afresh.  K enters through its Hermitian
part, so E is defined for every complex K and its K-gradient is Hermitian.  The backward is the
adjoint pullback of the six trajectories (the same integration as the Riemannian gradient,
src/observables.jl:161-212) with the coefficients of dR and dQ kept apart; it is computed in the
same launch as the energy, because the forward and adjoint trajectories are independent and run
concurrently.  The Gauss-node samples of every trajectory stay in device memory between the
integration and the pullback (about 0.6 GB at D = 64 with 256 nodes), so nothing is recomputed.
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import SolverFailure
from .observables import fg_phi4_rk
from .state import IntegratorParameters, State


def _host(x: torch.Tensor) -> np.ndarray:
    return x.detach().resolve_conj().to("cpu", torch.complex128).numpy()


class Phi4Energy(torch.autograd.Function):
    """E = Phi4Energy.apply(R, K, g, integrator, steady_tol, device); real float64 scalar."""

    @staticmethod
    def forward(ctx, R, K, g, integrator, steady_tol, device):
        want = bool(ctx.needs_input_grad[0] or ctx.needs_input_grad[1])
        k = _host(K)
        s = State(_host(R), (k + k.conj().T) / 2, steady_tol=steady_tol, device=device,
                  integrator_param=integrator)
        try:
            energy, GR, GK, stats = fg_phi4_rk(s, g, gradient=want)
        finally:
            s.close()
        ctx.stats = stats
        if want:
            ctx.save_for_backward(GR.to(R.device), GK.to(K.device))
        return torch.tensor(energy, dtype=torch.float64, device=R.device)

    @staticmethod
    def backward(ctx, grad_out):
        GR, GK = ctx.saved_tensors
        gr = GR * grad_out if ctx.needs_input_grad[0] else None
        gk = GK * grad_out if ctx.needs_input_grad[1] else None
        return gr, gk, None, None, None, None


def phi4_energy(R: torch.Tensor, K: torch.Tensor, g: float, *, integrator: IntegratorParameters | None = None,
                steady_tol: float = 1e-11, device: int = 0) -> torch.Tensor:
    """Differentiable fg_phi4 energy of State(R, K) at coupling g."""
    if R.dtype != torch.complex128 or K.dtype != torch.complex128:
        raise TypeError("R and K must be complex128")
    return Phi4Energy.apply(R, K, float(g), integrator or IntegratorParameters(), float(steady_tol), int(device))


def lbfgs_rk(R0, K0, g: float, *, n_iter: int = 50, history: int = 20, integrator=None, device: int = 0,
             gradtol: float = 1e-10, c1: float = 1e-4, max_backtracks: int = 40):
    """Euclidean L-BFGS in (R, K) with the gradient of Phi4Energy and an Armijo backtracking line
    search.  A trial whose stationary state or evaluation fails (SolverFailure) is rejected and the
    step halved, as the reference's line search does (src/optimization.jl:45-47).

    Returns (R, K, energies, evaluations): energies[i] is the energy after i accepted iterations."""
    dev = torch.device("cuda", device)
    x = torch.stack([torch.view_as_real(torch.as_tensor(np.asarray(R0), dtype=torch.complex128)),
                     torch.view_as_real(torch.as_tensor(np.asarray(K0), dtype=torch.complex128))]).to(dev)
    evaluations = 0

    def evaluate(y):
        nonlocal evaluations
        evaluations += 1
        R = torch.view_as_complex(y[0]).clone().requires_grad_(True)
        K = torch.view_as_complex(y[1]).clone().requires_grad_(True)
        e = phi4_energy(R, K, g, integrator=integrator, device=device)
        e.backward()
        return float(e), torch.stack([torch.view_as_real(R.grad), torch.view_as_real(K.grad)])

    f, gr = evaluate(x)
    energies = [f]
    pairs = []
    for it in range(n_iter):
        if float(gr.abs().max()) <= gradtol:
            break
        q = gr.clone()
        alphas = []
        for s_k, y_k, rho_k in reversed(pairs):
            a = rho_k * float((s_k * q).sum())
            q -= a * y_k
            alphas.append(a)
        if pairs:
            s_k, y_k, _ = pairs[-1]
            q *= float((s_k * y_k).sum() / (y_k * y_k).sum())
        else:
            q *= 0.01 / float(gr.abs().max())     # first step moves no entry by more than 0.01
        for (s_k, y_k, rho_k), a in zip(pairs, reversed(alphas)):
            b = rho_k * float((y_k * q).sum())
            q += (a - b) * s_k
        d = -q
        slope = float((gr * d).sum())
        if slope >= 0:                            # not a descent direction: restart from steepest descent
            pairs.clear()
            d = -gr * (0.01 / float(gr.abs().max()))
            slope = float((gr * d).sum())
        t = 1.0
        for _ in range(max_backtracks):
            trial = x + t * d
            try:
                f_new, g_new = evaluate(trial)
            except SolverFailure:
                t *= 0.5
                continue
            if f_new <= f + c1 * t * slope:
                break
            t *= 0.5
        else:
            raise SolverFailure(f"lbfgs_rk: no acceptable step at iteration {it}")
        s_k, y_k = trial - x, g_new - gr
        sy = float((s_k * y_k).sum())
        if sy > 1e-12 * float(s_k.norm() * y_k.norm()):
            pairs.append((s_k, y_k, 1.0 / sy))
            if len(pairs) > history:
                pairs.pop(0)
        x, f, gr = trial, f_new, g_new
        energies.append(f)
    R = torch.view_as_complex(x[0].contiguous()).cpu().numpy()
    K = torch.view_as_complex(x[1].contiguous()).cpu().numpy()
    return R, (K + K.conj().T) / 2, energies, evaluations
