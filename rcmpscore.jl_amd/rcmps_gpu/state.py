"""Left-canonical RCMPS states on the GPU (src/state.jl:1-77).

A State holds R, K (K Hermitian), Q = -iK - R^+R/2 and the stationary density rho0 (solved on the
GPU at construction, src/transfer.jl:74-115), plus the parameter records of the reference:
IntegratorParameters, ModelParameters and ManifoldParameters.  Matrices are complex128, row-major.
"""
from __future__ import annotations

import copy
import ctypes
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib


@dataclass
class IntegratorParameters:
    """`max_point` is the physical half-window (mass one); `n_quad` the Gauss nodes per half
    (src/state.jl:4-10)."""
    n_quad: int = 128
    max_point: float = 20.0
    atol: float = 1e-11
    rtol: float = 1e-10
    maxiters: int = 1_000_000


@dataclass
class ModelParameters:
    name: str = "shg"
    coupling: float = 1.0
    mass: float = 1.0


@dataclass
class ManifoldParameters:
    metric_reg: float = 1e-3
    precond_reg_prefactor: float = 1e-2
    precond_reg: float = 1e-4


def _as_matrix(x, name):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
    if a.ndim != 2 or a.shape[0] != a.shape[1] or a.shape[0] == 0:
        raise ValueError(f"{name} must be a nonempty square matrix")
    return a


class State:
    """State(R, K; rho=None, steady_tol=1e-11) with its stationary density (src/state.jl:40-53)."""

    def __init__(self, R, K, *, rho=None, steady_tol=1e-11, device: int = 0,
                 integrator_param: IntegratorParameters | None = None,
                 model_param: ModelParameters | None = None,
                 manifold_param: ManifoldParameters | None = None, info: str = ""):
        R = _as_matrix(R, "R")
        K = _as_matrix(K, "K")
        if K.shape != R.shape:
            raise ValueError("K and R must have equal dimensions")
        if not (np.isfinite(R).all() and np.isfinite(K).all()):
            raise ValueError("nonfinite state parameters")
        if np.linalg.norm(K - K.conj().T) > 1e-12 * max(1.0, np.linalg.norm(K)):
            raise ValueError("K must be Hermitian")
        self.device = int(device)
        self.D = R.shape[0]
        self.integrator_param = integrator_param or IntegratorParameters()
        self.model_param = model_param or ModelParameters()
        self.manifold_param = manifold_param or ManifoldParameters()
        self.e0 = math.nan
        self.grad = np.zeros_like(R)
        self.info = info
        self._handle = None
        lib = _lib.load()
        ctx = _lib.context(self.device)
        dev = torch.device("cuda", self.device)
        self._R = torch.from_numpy(R).to(dev)
        self._K = torch.from_numpy(K).to(dev)
        rho_t = None
        if rho is not None:
            rho_t = torch.from_numpy(_as_matrix(rho, "rho")).to(dev)
        torch.cuda.synchronize(dev)
        handle = ctypes.c_void_p()
        stats = _lib.Stats()
        st = lib.rcmps_state_create(ctx, self.D, self._R.data_ptr(), self._K.data_ptr(),
                                    rho_t.data_ptr() if rho_t is not None else None, float(steady_tol),
                                    ctypes.byref(handle), ctypes.byref(stats))
        _lib.check(st, "stationary state", stats)
        self._handle = handle
        _lib.register_state(self)
        self.steady_stats = stats.as_dict()
        q = torch.empty((self.D, self.D), dtype=torch.complex128, device=dev)
        r0 = torch.empty((self.D, self.D), dtype=torch.complex128, device=dev)
        _lib.check(lib.rcmps_state_get(ctx, handle, 0, q.data_ptr()), "state Q")
        _lib.check(lib.rcmps_state_get(ctx, handle, 1, r0.data_ptr()), "state rho")
        self._Q, self._rho = q, r0
        self.R = R
        self.K = (K + K.conj().T) / 2
        self.Q = q.cpu().numpy()
        self.rho = r0.cpu().numpy()

    # reference names
    @property
    def ρ0(self):
        return self.rho

    @property
    def handle(self):
        return self._handle

    @property
    def nojump_form(self):
        """("schur" | "eigen", kappa): the basis the no-jump preconditioner of this state applies in,
        and the condition estimate of Q's eigenvectors that chose it (rcmps_gpu.nojump_threshold)."""
        schur, kappa = ctypes.c_int(), ctypes.c_double()
        _lib.check(_lib.load().rcmps_state_nojump_form(self._handle, ctypes.byref(schur), ctypes.byref(kappa)),
                   "nojump form")
        return ("schur" if schur.value else "eigen"), kappa.value

    def close(self):
        if self._handle is not None and not _lib.closed():
            try:
                _lib.load().rcmps_state_destroy(self._handle)
            finally:
                self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def copy_param(self, src: "State") -> "State":
        """copy_param! (src/state.jl:72-77)."""
        self.model_param = copy.deepcopy(src.model_param)
        self.integrator_param = copy.deepcopy(src.integrator_param)
        self.manifold_param = copy.deepcopy(src.manifold_param)
        self.info = src.info
        return self


def _complex_normal(rng, d):
    return (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / math.sqrt(2)


def random_state_arrays(D: int, *, scaling=1.0, seed=1234, rng=None):
    """R, K of RandomState (src/state.jl:55-60): complex normal entries (variance 1), the matrix
    that becomes K drawn first, then R; K = scaling (k + k^+).  numpy PCG64 seeded with `seed`
    stands in for Julia's MersenneTwister, so draws differ from Julia's for the same seed."""
    if D <= 0:
        raise ValueError("D must be positive")
    rng = rng if rng is not None else np.random.default_rng(seed)
    k = _complex_normal(rng, D)
    r = scaling * _complex_normal(rng, D)
    return r, scaling * (k + k.conj().T)


def RandomState(D: int, *, scaling=1.0, seed=1234, rng=None, **kw) -> State:
    r, k = random_state_arrays(D, scaling=scaling, seed=seed, rng=rng)
    return State(r, k, **kw)


def random_state_z2_arrays(D: int, *, scaling=1.0, seed=1234, rng=None):
    """RandomStateZ2 (src/state.jl:61-70): R odd, K even under diag(1,...,1,-1,...,-1)."""
    if D <= 0 or D % 2:
        raise ValueError("Z2 states require a positive even D")
    rng = rng if rng is not None else np.random.default_rng(seed)
    d = D // 2
    k1, k2 = _complex_normal(rng, d), _complex_normal(rng, d)
    z = np.zeros((d, d), np.complex128)
    r = np.block([[z, _complex_normal(rng, d)], [_complex_normal(rng, d), z]])
    k = np.block([[k1 + k1.conj().T, z], [z, k2 + k2.conj().T]])
    return scaling * r, scaling * k


def RandomStateZ2(D: int, *, scaling=1.0, seed=1234, rng=None, **kw) -> State:
    r, k = random_state_z2_arrays(D, scaling=scaling, seed=seed, rng=rng)
    return State(r, k, **kw)
