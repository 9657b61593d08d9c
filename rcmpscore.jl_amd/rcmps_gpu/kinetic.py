"""Kinetic terms on the infinite line (kinetic_backend="krylov"): the angular resolvent
(2/pi) int_0^{pi/2} tr[C (1 - cos(theta) L)^{-1} (rho0 C^+)] dtheta of src/kinetic.jl:25-124, with
the shifted solves of every angular node taken from one matrix-free Arnoldi basis per operator
and direction (src/kinetic_krylov.jl), all on the GPU (librcmps_hip.so).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .state import State


@dataclass
class KineticOptions:
    """Angular quadrature controls (src/kinetic.jl:9-16)."""
    minnodes: int = 32
    maxnodes: int = 512
    atol: float = 1e-11
    rtol: float = 1e-10


@dataclass
class KineticKrylovOptions:
    """Matrix-free resolvent controls (src/kinetic_krylov.jl:8-13); fallback_tolerance defaults to
    tolerance (strict acceptance)."""
    tolerance: float = 1e-12
    fallback_tolerance: float | None = None
    initialdim: int = 16
    maxdim: int = 512


def apply_options(device: int, kinetic_options: KineticOptions | None, krylov_options: KineticKrylovOptions | None):
    k = kinetic_options or KineticOptions()
    q = krylov_options or KineticKrylovOptions()
    fb = q.tolerance if q.fallback_tolerance is None else q.fallback_tolerance
    opts = _lib.KineticOptions(int(k.minnodes), int(k.maxnodes), float(k.atol), float(k.rtol), float(q.tolerance),
                               float(fb), int(q.initialdim), int(q.maxdim))
    _lib.check(_lib.load().rcmps_set_kinetic_options(_lib.context(device), ctypes.byref(opts)), "kinetic options")


def _kinetic(s: State, which: int, gradient: bool, kinetic_backend: str, kinetic_options, krylov_options,
             return_stats: bool):
    if kinetic_backend != "krylov":
        raise NotImplementedError("only the matrix-free (krylov) infinite-line kinetic backend is built")
    apply_options(s.device, kinetic_options, krylov_options)
    lib = _lib.load()
    ctx = _lib.context(s.device)
    dev = torch.device("cuda", s.device)
    G = torch.empty((s.D, s.D), dtype=torch.complex128, device=dev) if gradient else None
    v = ctypes.c_double()
    stats = _lib.Stats()
    st = lib.rcmps_kinetic(ctx, s.handle, which, float(s.manifold_param.metric_reg), ctypes.byref(v), None, None,
                           G.data_ptr() if gradient else None, ctypes.byref(stats))
    _lib.check(st, "kinetic", stats)
    out = (v.value, G.cpu().numpy()) if gradient else v.value
    if return_stats:
        return out, stats.as_dict()
    return out


def kinetic1(s: State, *, gradient: bool = True, kinetic_backend: str = "krylov", kinetic_options=None,
             krylov_options=None, return_stats: bool = False):
    """kinetic1 (src/observables.jl:219) on the infinite line: value, or (value, Riemannian gradient)."""
    return _kinetic(s, 1, gradient, kinetic_backend, kinetic_options, krylov_options, return_stats)


def kinetic2(s: State, *, gradient: bool = True, kinetic_backend: str = "krylov", kinetic_options=None,
             krylov_options=None, return_stats: bool = False):
    """kinetic2 (src/observables.jl:220), the [Q, R] term, on the infinite line."""
    return _kinetic(s, 2, gradient, kinetic_backend, kinetic_options, krylov_options, return_stats)
