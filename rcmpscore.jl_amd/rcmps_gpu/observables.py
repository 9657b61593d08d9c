"""fg_phi4: energy and Riemannian gradient of the phi^4 model (src/observables.jl:226-257).

The kinetic terms and the normal-ordered quartic are all evaluated as finite-window ODE
observables (kinetic_backend=:ode, the Julia default): six trajectories (three terms, forward
and adjoint) integrated by the persistent HIP DOP853 kernel, the pullback over the Gauss nodes,
the boundary Poisson correction and the metric solve, all in librcmps_hip.so.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .state import IntegratorParameters, State

PARALLEL = (None, "serial", "terms", "trajectories", "both")
PRECISIONS = ("fp64", "split")


KINETIC_BACKENDS = ("ode", "krylov")


def _flags(boundary: bool, precision: str, kinetic_backend: str = "ode") -> int:
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {PRECISIONS}")
    if kinetic_backend not in KINETIC_BACKENDS:
        if kinetic_backend == "resolvent":
            raise NotImplementedError("the dense Schur kinetic backend is not built; use kinetic_backend='krylov'")
        raise ValueError("unknown kinetic backend")
    return ((0 if boundary else _lib.FLAG_NO_BOUNDARY) | (_lib.FLAG_SPLIT if precision == "split" else 0)
            | (_lib.FLAG_KINETIC_KRYLOV if kinetic_backend == "krylov" else 0))


def _integrator(p: IntegratorParameters, initial_step: float = 0.01) -> _lib.Integrator:
    if p.n_quad < 8 or not (p.max_point > 0 and p.atol > 0 and p.rtol > 0 and p.maxiters > 0):
        raise ValueError("invalid integration controls")
    return _lib.Integrator(int(p.n_quad), float(p.max_point), float(p.atol), float(p.rtol),
                           int(p.maxiters), float(initial_step))


def fg_phi4_raw(s: State, g: float, *, boundary: bool = True, integrator: IntegratorParameters | None = None,
                metric_reg: float | None = None, precision: str = "fp64", kinetic_backend: str = "ode",
                kinetic_options=None, krylov_options=None, phase_timers: bool = False, z2_blocks: bool = True):
    """(energy, gradient, raw covector, raw environment, stats dict) on device tensors.

    precision="split" runs the trajectory products on the INT8 MFMA over exact 8-bit slices of
    the FP64 operands; everything else stays FP64.  phase_timers=True fills stats["phase_us"]
    (the trajectory kernel's per-phase timers, which otherwise read no clock).  z2_blocks=False makes
    the trajectory products ignore a Z2 state's zero parity blocks (the same bits, every k-tile)."""
    lib = _lib.load()
    ctx = _lib.context(s.device)
    ip = _integrator(integrator or s.integrator_param)
    reg = s.manifold_param.metric_reg if metric_reg is None else metric_reg
    if reg < 0:
        raise ValueError("metric regularizer must be nonnegative")
    dev = torch.device("cuda", s.device)
    d = s.D
    G = torch.empty((d, d), dtype=torch.complex128, device=dev)
    cov = torch.empty_like(G)
    env = torch.empty_like(G)
    e = ctypes.c_double()
    stats = _lib.Stats()
    flags = (_flags(boundary, precision, kinetic_backend) | (_lib.FLAG_PHASE_TIMERS if phase_timers else 0)
             | (0 if z2_blocks else _lib.FLAG_DENSE))
    if kinetic_backend == "krylov":
        from .kinetic import apply_options
        apply_options(s.device, kinetic_options, krylov_options)
    st = lib.rcmps_fg_phi4(ctx, s.handle, float(g), ctypes.byref(ip), float(reg), flags, ctypes.byref(e),
                           G.data_ptr(), cov.data_ptr(), env.data_ptr(), ctypes.byref(stats))
    _lib.check(st, "fg_phi4", stats)
    return e.value, G, cov, env, stats.as_dict()


def fg_phi4(s: State, g: float, *, gradient: bool = True, kinetic_backend: str = "ode", parallel=None,
            boundary: bool = True, return_stats: bool = False, precision: str = "fp64", kinetic_options=None,
            krylov_options=None, phase_timers: bool = False, z2_blocks: bool = True):
    """fg_phi4(s, g) -> (energy, gradient) with dE[W] = metric(s, gradient, W) for horizontal W.

    gradient=False returns the energy only.  `parallel` is accepted for signature compatibility:
    the six trajectories always run concurrently in one launch."""
    if parallel not in PARALLEL:
        raise ValueError("parallel must be :serial, :terms, :trajectories, or :both")
    energy, G, _, _, stats = fg_phi4_raw(s, g, boundary=boundary, precision=precision, kinetic_backend=kinetic_backend,
                                         kinetic_options=kinetic_options, krylov_options=krylov_options,
                                         phase_timers=phase_timers, z2_blocks=z2_blocks)
    if not gradient:
        return (energy, stats) if return_stats else energy
    grad = G.cpu().numpy()
    return (energy, grad, stats) if return_stats else (energy, grad)


def fg_phi4_batch(states, couplings, *, gradient: bool = True, precision: str = "fp64",
                  integrator: IntegratorParameters | None = None, metric_reg: float | None = None,
                  return_stats: bool = False):
    """fg_phi4 of many independent states (any mix of D and g) in one call: the trajectories of
    states of one padded size run in one persistent launch.  Returns (energies, gradients) with
    gradients as numpy arrays (None when gradient=False), plus the per-state stats dicts when
    return_stats.  Raises SolverFailure if any evaluation failed."""
    states = list(states)
    gs = [float(x) for x in couplings]
    if not states or len(gs) != len(states):
        raise ValueError("one coupling per state is required")
    device = states[0].device
    if any(s.device != device for s in states):
        raise ValueError("all states must live on one device")
    lib = _lib.load()
    ctx = _lib.context(device)
    ip = _integrator(integrator or states[0].integrator_param)
    reg = states[0].manifold_param.metric_reg if metric_reg is None else metric_reg
    dev = torch.device("cuda", device)
    n = len(states)
    grads = [torch.empty((s.D, s.D), dtype=torch.complex128, device=dev) for s in states] if gradient else None
    handles = (ctypes.c_void_p * n)(*[s.handle.value for s in states])
    garr = (ctypes.c_void_p * n)(*[g.data_ptr() for g in grads]) if gradient else None
    energies = (ctypes.c_double * n)()
    stats = (_lib.Stats * n)()
    st = lib.rcmps_fg_phi4_batch(ctx, n, handles, (ctypes.c_double * n)(*gs), ctypes.byref(ip), float(reg),
                                 _flags(True, precision), energies, garr, stats)
    _lib.check(st, "fg_phi4_batch", stats[0] if st == 0 else next((x for x in stats if x.status), stats[0]))
    out_e = list(energies)
    out_g = [g.cpu().numpy() for g in grads] if gradient else None
    if return_stats:
        return out_e, out_g, [x.as_dict() for x in stats]
    return out_e, out_g


def fg_phi4_rk(s: State, g: float, *, gradient: bool = True, boundary: bool = True,
               integrator: IntegratorParameters | None = None, precision: str = "fp64"):
    """(energy, grad_R, grad_K, stats): the energy of State(R, K) and its Euclidean gradient in
    (R, K) as device tensors, dE = Re sum(conj(grad_R) dR + conj(grad_K) dK) for Hermitian dK.
    gradient=False integrates the forward trajectories only and returns None gradients."""
    lib = _lib.load()
    ctx = _lib.context(s.device)
    ip = _integrator(integrator or s.integrator_param)
    dev = torch.device("cuda", s.device)
    GR = GK = None
    if gradient:
        GR = torch.empty((s.D, s.D), dtype=torch.complex128, device=dev)
        GK = torch.empty_like(GR)
    e = ctypes.c_double()
    stats = _lib.Stats()
    flags = _flags(boundary, precision)
    st = lib.rcmps_fg_phi4_rk(ctx, s.handle, float(g), ctypes.byref(ip), flags, ctypes.byref(e),
                              GR.data_ptr() if gradient else None, GK.data_ptr() if gradient else None,
                              ctypes.byref(stats))
    _lib.check(st, "fg_phi4_rk", stats)
    return e.value, GR, GK, stats.as_dict()


def boundary_environment(s: State, a, *, tol: float = 1e-11, no_alloc: bool = False) -> np.ndarray:
    """Solve (-L* + |I><rho|) H = a - tr(rho a) I (src/transfer.jl:122-153).

    no_alloc=True runs the same solve through rcmps_boundary_environment_noalloc, whose device
    workspace is a fixed library-owned array (D <= 64; larger states raise ValueError)."""
    lib = _lib.load()
    ctx = _lib.context(s.device)
    dev = torch.device("cuda", s.device)
    at = torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.complex128))).to(dev)
    torch.cuda.synchronize(dev)
    h = torch.empty_like(at)
    stats = _lib.Stats()
    fn = lib.rcmps_boundary_environment_noalloc if no_alloc else lib.rcmps_boundary_environment
    _lib.check(fn(ctx, s.handle, at.data_ptr(), float(tol), h.data_ptr(), ctypes.byref(stats)),
               "boundary_environment", stats)
    return h.cpu().numpy()


def trajectory_samples(s: State, system: int) -> np.ndarray:
    """Dense-output samples of one system of the last fg_phi4 on this device:
    (2 n_quad + 1, live channels, D, D)."""
    lib = _lib.load()
    ctx = _lib.context(s.device)
    nlive = (1, 2, 1, 2, 3, 4)[system]
    npts = 2 * s.integrator_param.n_quad + 1
    out = torch.empty((npts, nlive, s.D, s.D), dtype=torch.complex128, device=torch.device("cuda", s.device))
    _lib.check(lib.rcmps_trajectory_samples(ctx, int(system), out.data_ptr(), out.numel() * 2),
               "trajectory_samples")
    return out.cpu().numpy()
