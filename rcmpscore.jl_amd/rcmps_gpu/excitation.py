"""Tangent-space excitations on the GPU: the Hamiltonian action, white sector coordinates and the
preconditioned Davidson eigensolver (src/excitation_environments.jl, src/excitations.jl,
src/excitation_action.jl, src/davidson.jl, src/excitation_eigsolve.jl).

The action F(W) (rcmps_excitation_action) runs on the device for whole batches of tangents; the
white coordinates, the Davidson iteration and the kinetic preconditioner's dense eigenproblem act
on sector vectors of length D^2/2 on the host, as in the reference.

Implemented: ExcitationModel (:phi4), ExcitationContext, excitation_hamiltonian_action,
excitation_poisson / action_dual_poisson, TangentCoordinates, DavidsonOptions, davidson,
ExcitationOperator, kinetic_preconditioner, field_seed, excitation_eigsolve,
iterative_dispersion, dispersion_relation (iterative), relativistic_diagnostics, excitation_element,
tangent_metric.
Validation computes the reference's four energies: fresh actions at the standard and refined
settings, and the bilinear (connected-element) energies at both (excitation_element on the device).
"""
from __future__ import annotations

import copy
import ctypes
import math
import time
from dataclasses import dataclass, field, replace

import numpy as np
import torch

from . import _lib
from ._lib import SolverFailure
from .kinetic import KineticKrylovOptions, KineticOptions, apply_options
from .state import IntegratorParameters, State


# --------------------------------------------------------------------------------------------
# model and context
# --------------------------------------------------------------------------------------------
@dataclass(frozen=True)
class ExcitationModel:
    """The phi^4 Hamiltonian from the ground-state terms (src/excitations.jl:1-24): kinetic terms
    with weights (1, 1) and g :phi^4:."""
    name: str = "phi4"
    coupling: float = 1.0

    def __post_init__(self):
        if self.name != "phi4":
            raise NotImplementedError(f"excitation model {self.name!r}: only phi4 is built")
        if not math.isfinite(self.coupling):
            raise ValueError("nonfinite model parameter")

    @property
    def kinetic(self):
        return (1.0, 1.0)

    @staticmethod
    def of(s: State) -> "ExcitationModel":
        if s.model_param.name != "phi4":
            raise NotImplementedError(f"excitation model {s.model_param.name!r}: only phi4 is built")
        return ExcitationModel("phi4", float(np.real(s.model_param.coupling)))


def check_kinetic_nodes(n):
    if not (isinstance(n, (int, np.integer)) and n >= 2):
        raise ValueError("kinetic_nodes must be an integer of at least 2")
    return int(n)


# Largest D for which the excitation kinetic resolvents use the transfer's dense eigendecomposition
# (one host eigensolve of the D^2 x D^2 transfer per state: about 1.5 s at D = 32, 16 MB per factor).
DENSE_TRANSFER_MAX_D = 32
# eigenvector condition estimate above which the dense route is not used (Krylov bases instead)
DENSE_TRANSFER_MAX_COND = 1e10


def transfer_eigen(s: State, enable: bool = True):
    """Attach (or, with enable=False, remove) the eigendecomposition of the transfer superoperator
    L(X) = Q X + X Q^+ + R X R^+ to the state on the device, so that the excitation kinetic resolvents
    of every momentum are products with its eigenvectors (the shifts L - ip and L* + ip keep them).
    Returns the condition estimate |W|_F |W^-1|_F, or None when the dense route is not used (D above
    DENSE_TRANSFER_MAX_D, or an ill-conditioned eigenbasis)."""
    lib = _lib.load()
    ctx = _lib.context(s.device)
    if not enable or s.D > DENSE_TRANSFER_MAX_D:
        _lib.check(lib.rcmps_state_set_transfer_eigen(ctx, s.handle, 0, None, None, None), "transfer eigen")
        s._transfer_cond = None
        s._transfer_unusable = enable
        return None
    d = s.D
    eye = np.eye(d)
    L = np.kron(s.Q, eye) + np.kron(eye, s.Q.conj()) + np.kron(s.R, s.R.conj())
    mu, W = np.linalg.eig(L)
    Winv = np.linalg.inv(W)
    cond = float(np.linalg.norm(W) * np.linalg.norm(Winv))
    if not (np.isfinite(cond) and cond <= DENSE_TRANSFER_MAX_COND):
        _lib.check(lib.rcmps_state_set_transfer_eigen(ctx, s.handle, 0, None, None, None), "transfer eigen")
        s._transfer_cond = None
        s._transfer_unusable = True
        return None
    dev = _device(s)
    wt = torch.from_numpy(np.ascontiguousarray(W.T)).to(dev)        # column-major W
    wit = torch.from_numpy(np.ascontiguousarray(Winv.T)).to(dev)
    mut = torch.from_numpy(np.ascontiguousarray(mu)).to(dev)
    torch.cuda.synchronize(dev)
    _lib.check(lib.rcmps_state_set_transfer_eigen(ctx, s.handle, d * d, wt.data_ptr(), wit.data_ptr(), mut.data_ptr()),
               "transfer eigen")
    s._transfer_cond = cond
    s._transfer_unusable = False
    return cond


class ExcitationContext:
    """ExcitationContext(s, p; tol) (src/excitation_environments.jl:18-40) on the device.  Built
    matrix-free: the mixing check of the dense reference is replaced by the residual checks of the
    Poisson and resolvent solves.  `integrator` overrides the state's integration settings for the
    potential term (the refined settings of the validation).  dense_transfer: for D <= 32 the state's
    transfer eigendecomposition (transfer_eigen, once per state) serves the kinetic resolvents;
    False keeps the Krylov bases (and removes a decomposition the state carries)."""

    def __init__(self, s: State, p: float = 0.0, *, tol: float = 1e-10, integrator=None, dense_transfer: bool = True):
        if not (math.isfinite(p) and math.isfinite(tol) and tol > 0):
            raise ValueError("invalid momentum or environment tolerance")
        self.state = s
        self.momentum = float(p)
        self.tol = float(tol)
        self.integrator = copy.deepcopy(integrator or s.integrator_param)
        lib = _lib.load()
        if not dense_transfer:
            if getattr(s, "_transfer_cond", None) is not None:
                transfer_eigen(s, False)
        elif getattr(s, "_transfer_cond", None) is None and not getattr(s, "_transfer_unusable", False):
            transfer_eigen(s, True)
        handle = ctypes.c_void_p()
        _lib.check(lib.rcmps_excitation_create(_lib.context(s.device), s.handle, self.momentum, self.tol,
                                               ctypes.byref(handle)), "excitation context")
        self._handle = handle
        _lib.register_state(self)

    def with_settings(self, *, momentum=None, integrator=None) -> "ExcitationContext":
        """ExcitationContext(base; momentum, integrator) (src/excitation_environments.jl:58-71)."""
        return ExcitationContext(self.state, self.momentum if momentum is None else momentum, tol=self.tol,
                                 integrator=integrator or self.integrator)

    @property
    def handle(self):
        return self._handle

    def close(self):
        if getattr(self, "_handle", None) is not None and not _lib.closed():
            try:
                _lib.load().rcmps_excitation_destroy(self._handle)
            finally:
                self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _device(s: State):
    return torch.device("cuda", s.device)


def _to_dev(x, s: State):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.complex128))).to(_device(s))


def excitation_poisson(ctx: ExcitationContext, b, shift: int = 0):
    """(-L + i shift p + |rho><I|) x = b - tr(b) rho, trace-projected (src/excitation_environments.jl:75-85)."""
    return _poisson(ctx, b, False, shift)


def action_dual_poisson(ctx: ExcitationContext, b, shifted: bool = False):
    """(-L* - i p [shifted] + |I><rho|) x = b - tr(rho b) I, projected (src/excitation_action.jl:85-94)."""
    return _poisson(ctx, b, True, 1 if shifted else 0)


def _poisson(ctx, b, adjoint, shift):
    s = ctx.state
    bd = _to_dev(b, s)
    xd = torch.empty_like(bd)
    torch.cuda.synchronize(_device(s))
    st = _lib.load().rcmps_excitation_poisson(_lib.context(s.device), ctx.handle, bd.data_ptr(), int(adjoint),
                                              int(shift), xd.data_ptr())
    _lib.check(st, "excitation Poisson solve")
    return xd.cpu().numpy()


def _integrator(p: IntegratorParameters):
    return _lib.Integrator(int(p.n_quad), float(p.max_point), float(p.atol), float(p.rtol), int(p.maxiters), 0.01)


# Krylov controls of the theta-resolvent solves: the action's (true relative residual 1e-12), and
# the coarse kinetic matrix of the preconditioner's (1e-9: a preconditioner needs no more)
ACTION_KRYLOV = KineticKrylovOptions()
PRECONDITIONER_KRYLOV = KineticKrylovOptions(tolerance=1e-9, fallback_tolerance=1e-9)


def action_batch(ctx, ws, model: ExcitationModel, kinetic_nodes: int, parts: int = 0, return_stats: bool = False,
                 krylov_options: KineticKrylovOptions | None = None, z2_blocks: bool = True,
                 persistent: bool | None = None):
    """F(W) for a list of D x D tangents in one device call (every kinetic Krylov iteration and ODE
    step shared); kinetic part on the fixed grid of `kinetic_nodes` theta nodes.  `ctx` is one
    ExcitationContext or one per tangent (momenta may differ; state and integration settings must
    not).  z2_blocks=False makes the action ODE ignore a Z2 state's zero parity blocks (same bits);
    persistent=True integrates it in one persistent launch where every tile's workgroup fits,
    False by replayed launch graphs, None takes the faster for the size (same bits)."""
    ws = [np.asarray(w, dtype=np.complex128) for w in ws]
    ctxs = list(ctx) if isinstance(ctx, (list, tuple)) else [ctx] * len(ws)
    if len(ctxs) != len(ws):
        raise ValueError("one context per tangent is required")
    if not ws:
        return ([], {}) if return_stats else []
    s = ctxs[0].state
    ip0 = ctxs[0].integrator
    for c in ctxs:
        if c.state is not s:
            raise ValueError("all contexts of one call must share the state")
        if c.integrator != ip0:
            raise ValueError("all contexts of one call must share the integration settings")
    for w in ws:
        if w.shape != (s.D, s.D):
            raise ValueError("tangent dimensions do not match the state")
        if not np.isfinite(w).all():
            raise ValueError("nonfinite tangent")
    dev = _device(s)
    W = torch.from_numpy(np.ascontiguousarray(np.stack(ws))).to(dev)
    F = torch.empty_like(W)
    n = len(ws)
    wp = (ctypes.c_void_p * n)(*[W[i].data_ptr() for i in range(n)])
    fp = (ctypes.c_void_p * n)(*[F[i].data_ptr() for i in range(n)])
    hp = (ctypes.c_void_p * n)(*[c.handle for c in ctxs])
    kc = (ctypes.c_double * 2)(*model.kinetic)
    ip = _integrator(ip0)
    stats = _lib.Stats()
    apply_options(s.device, None, krylov_options or ACTION_KRYLOV)
    torch.cuda.synchronize(dev)
    st = _lib.load().rcmps_excitation_action_multi(_lib.context(s.device), n, hp, wp, fp, float(model.coupling), kc,
                                                   ctypes.byref(ip), int(kinetic_nodes),
                                                   int(parts) | (0 if z2_blocks else _lib.EXC_DENSE)
                                                   | {None: 0, True: _lib.EXC_PERSISTENT,
                                                      False: _lib.EXC_GRAPH}[persistent],
                                                   ctypes.byref(stats))
    if st == 10:
        phase = {20: "ket Poisson solve", 21: "kinetic forward Krylov", 22: "kinetic adjoint Krylov",
                 23: "boundary Poisson solve", 24: "action ODE"}.get(stats.failed_system, "?")
        raise SolverFailure(f"excitation action: solver failure in the {phase} (residual {stats.krylov_residual:.3e}, "
                            f"Krylov dimension {stats.kinetic_basis[2]})")
    _lib.check(st, "excitation action", stats)
    out = list(F.cpu().numpy())
    return (out, stats.as_dict()) if return_stats else out


def excitation_element(ctx, us, ws, *, model: ExcitationModel | None = None, kinetic_nodes: int = 256, parts: int = 0,
                       return_stats: bool = False):
    """Connected matrix elements <Phi_p(u)|H - E0|Phi_p(w)> per 2 pi delta(0) for lists of tangent pairs
    (excitation_element(ExcitationPair(ctx, u, w), model), src/excitations.jl:98-113) in one device
    call; `ctx` is one context or one per pair.  The kinetic part uses the fixed theta grid."""
    us = [np.asarray(u, dtype=np.complex128) for u in us]
    ws = [np.asarray(w, dtype=np.complex128) for w in ws]
    if len(us) != len(ws):
        raise ValueError("one u per w is required")
    ctxs = list(ctx) if isinstance(ctx, (list, tuple)) else [ctx] * len(ws)
    if len(ctxs) != len(ws):
        raise ValueError("one context per pair is required")
    if not ws:
        return np.zeros(0, complex)
    s = ctxs[0].state
    model = model or ExcitationModel.of(s)
    ip0 = ctxs[0].integrator
    for c in ctxs:
        if c.state is not s or c.integrator != ip0:
            raise ValueError("all contexts of one call must share the state and the integration settings")
    for x in us + ws:
        if x.shape != (s.D, s.D):
            raise ValueError("tangent dimensions do not match the state")
        if not np.isfinite(x).all():
            raise ValueError("nonfinite tangent")
    dev = _device(s)
    Ut = torch.from_numpy(np.ascontiguousarray(np.stack(us))).to(dev)
    Wt = torch.from_numpy(np.ascontiguousarray(np.stack(ws))).to(dev)
    n = len(ws)
    up = (ctypes.c_void_p * n)(*[Ut[i].data_ptr() for i in range(n)])
    wp = (ctypes.c_void_p * n)(*[Wt[i].data_ptr() for i in range(n)])
    hp = (ctypes.c_void_p * n)(*[c.handle for c in ctxs])
    kc = (ctypes.c_double * 2)(*model.kinetic)
    ip = _integrator(ip0)
    out = (ctypes.c_double * (2 * n))()
    stats = _lib.Stats()
    apply_options(s.device, None, ACTION_KRYLOV)
    torch.cuda.synchronize(dev)
    st = _lib.load().rcmps_excitation_element(_lib.context(s.device), n, hp, up, wp, float(model.coupling), kc,
                                              ctypes.byref(ip), int(check_kinetic_nodes(kinetic_nodes)), int(parts), out,
                                              ctypes.byref(stats))
    _lib.check(st, "excitation matrix element", stats)
    vals = np.array(out[0::2]) + 1j * np.array(out[1::2])
    return (vals, stats.as_dict()) if return_stats else vals


def tangent_metric(s: State, u, w):
    """The physical tangent inner product tr(w rho u^+) (src/excitation_environments.jl:88-92)."""
    return complex(np.trace(np.asarray(w) @ s.rho @ np.asarray(u).conj().T))


def excitation_hamiltonian_action(ctx: ExcitationContext, w, *, model: ExcitationModel | None = None,
                                  kinetic_nodes=None, kinetic_options: KineticOptions | None = None,
                                  parts: int = 0):
    """excitation_hamiltonian_action(ctx, W) (src/excitation_action.jl:194-228): F with
    <U, F(W)> = <Phi_p(U)|H - E0|Phi_p(W)>.  kinetic_nodes=None refines the theta grid by node
    doubling for this W (src/excitation_action.jl:180-192); an integer fixes it.  `w` may be a list
    of tangents (one batched device call; a list of actions is returned), and `ctx` a list of
    contexts, one per tangent (momenta may differ)."""
    s0 = (ctx[0] if isinstance(ctx, (list, tuple)) else ctx).state
    model = model or ExcitationModel.of(s0)
    single = not isinstance(w, (list, tuple))
    ws = [w] if single else list(w)
    if kinetic_nodes is not None:
        out = action_batch(ctx, ws, model, check_kinetic_nodes(kinetic_nodes), parts)
        return out[0] if single else out
    o = kinetic_options or KineticOptions()
    if not (o.minnodes >= 2 and o.maxnodes >= 2 * o.minnodes and o.atol > 0 and o.rtol >= 0):
        raise ValueError("invalid kinetic quadrature controls")
    pot = ([np.zeros((s0.D, s0.D), complex)] * len(ws) if parts == _lib.EXC_KINETIC else
           action_batch(ctx, ws, model, 2, _lib.EXC_POTENTIAL))
    if parts == _lib.EXC_POTENTIAL:
        return pot[0] if single else pot
    n = o.minnodes
    prev = action_batch(ctx, ws, model, n, _lib.EXC_KINETIC)
    while 2 * n <= o.maxnodes:
        n *= 2
        cur = action_batch(ctx, ws, model, n, _lib.EXC_KINETIC)
        if all(np.linalg.norm(c - p) <= o.atol + o.rtol * np.linalg.norm(c) for c, p in zip(cur, prev)):
            out = [c + q for c, q in zip(cur, pot)]
            return out[0] if single else out
        prev = cur
    raise SolverFailure(f"excitation kinetic action quadrature did not converge by {n} nodes")


# --------------------------------------------------------------------------------------------
# white coordinates (src/excitation_eigsolve.jl:1-130)
# --------------------------------------------------------------------------------------------
def _parity_error(s: State, z):
    zz = np.outer(z, z)
    return max(np.linalg.norm(zz * s.R + s.R) / max(1, np.linalg.norm(s.R)),
               np.linalg.norm(zz * s.K - s.K) / max(1, np.linalg.norm(s.K)), np.linalg.norm(zz * s.rho - s.rho))


def excitation_parity(s: State, parity="auto", tol=1e-12):
    """The diagonal Z2 parity of s as a +-1 vector, or None (src/excitation_eigsolve.jl:13-35)."""
    if parity is None:
        return None
    if isinstance(parity, str):
        if parity != "auto":
            raise ValueError("parity must be 'auto', None, or a +-1 vector")
        if s.D % 2:
            return None
        z = np.concatenate([np.ones(s.D // 2), -np.ones(s.D // 2)])
        return z if _parity_error(s, z) <= tol else None
    z = np.asarray(parity, dtype=float)
    if not (len(z) == s.D and set(np.unique(z)) == {-1.0, 1.0}):
        raise ValueError("parity must be a +-1 vector of length D with both signs")
    if _parity_error(s, z) > tol:
        raise ValueError("state is not Z2 symmetric under the supplied parity")
    return z


class TangentCoordinates:
    """White coordinates of one sector: c -> W = X B^+ with X supported on `indices` and
    B^+ rho B = 1 on the resolved metric support, so tr(W rho W^+) = |c|^2
    (src/excitation_eigsolve.jl:37-112).  gauge='density' rotates into the eigenbasis of rho per
    parity block; `state` is the state in the coordinates' gauge (a device State)."""

    def __init__(self, s: State, *, sector="auto", parity="auto", gauge="density", metric_tol=0.0, parity_tol=1e-12,
                 steady_tol=1e-11):
        if not (math.isfinite(metric_tol) and 0 <= metric_tol < 1):
            raise ValueError("metric_tol must lie in [0, 1)")
        if gauge not in ("density", "original"):
            raise ValueError("gauge must be 'density' or 'original'")
        z = excitation_parity(s, parity, parity_tol)
        if sector == "auto":
            sector = "all" if z is None else "odd"
        if sector not in ("all", "odd", "even"):
            raise ValueError("sector must be 'auto', 'all', 'odd' or 'even'")
        if sector != "all" and z is None:
            raise ValueError(f"the {sector} sector requires a Z2-symmetric state; supply its parity")
        d = s.D
        blocks = [np.arange(d)] if z is None else [np.where(z == 1)[0], np.where(z == -1)[0]]
        u = np.zeros((d, d), complex)
        lam = np.zeros(d)
        for b in blocks:
            vals, vecs = np.linalg.eigh(s.rho[np.ix_(b, b)])
            u[np.ix_(b, b)] = vecs
            lam[b] = vals
        cutoff = max(metric_tol, 10 * d * np.finfo(float).eps) * lam.max()
        keep = lam > cutoff
        if not keep.any():
            raise SolverFailure("empty tangent metric support")
        if gauge == "density":
            rho = u.conj().T @ s.rho @ u
            t = State(u.conj().T @ s.R @ u, u.conj().T @ s.K @ u, rho=(rho + rho.conj().T) / 2, steady_tol=steady_tol,
                      device=s.device)
            t.copy_param(s)
            t.e0 = s.e0
            rotation = u
        else:
            t, rotation = s, np.eye(d, dtype=complex)
        B = np.zeros((d, d), complex)
        colkeep = np.ones(d, bool)
        for blk in blocks:
            ix = np.ix_(blk, blk)
            done = False
            if keep.all():
                try:
                    L = np.linalg.cholesky(t.rho[ix])
                    B[ix] = np.linalg.inv(L.conj().T)
                    done = True
                except np.linalg.LinAlgError:
                    pass
            if not done:
                vals, vecs = np.linalg.eigh(t.rho[ix]) if gauge == "density" else (lam[blk], u[ix])
                mask = vals > cutoff
                B[ix] = vecs @ np.diag(mask / np.sqrt(np.maximum(vals, np.finfo(float).tiny)))
                colkeep[blk] = mask
        columns = np.where(colkeep)[0]
        allowed = (lambda i, j: True) if sector == "all" else (lambda i, j: (sector == "odd") == (z[i] == z[j]))
        self.indices = [(i, j) for j in columns for i in range(d) if allowed(i, j)]
        if not self.indices:
            raise SolverFailure("empty tangent sector")
        self._rows = np.array([i for i, _ in self.indices])
        self._cols = np.array([j for _, j in self.indices])
        defect = np.linalg.norm(B.conj().T @ t.rho @ B - np.diag(np.isin(np.arange(d), columns).astype(float)))
        if defect > 1e-6:
            raise SolverFailure(f"tangent whitening is unresolved: metric defect {defect}")
        self.state, self.rotation, self.whitening = t, rotation, B
        self.parity, self.sector, self.gauge = z, sector, gauge
        self.parity_error = math.nan if z is None else _parity_error(s, z)
        self.metric_rank = int(keep.sum())
        self.discarded_metric_weight = float(np.abs(lam[~keep]).sum())
        self.metric_defect = float(defect)
        self.density_condition = float(lam.max() / lam[keep].min())

    def __len__(self):
        return len(self.indices)

    def tangent_matrix(self, x):
        m = np.zeros((self.state.D, self.state.D), complex)
        m[self._rows, self._cols] = x
        return m @ self.whitening.conj().T

    def tangent_coordinates(self, w):
        return (np.asarray(w) @ self.state.rho @ self.whitening)[self._rows, self._cols]

    def covector_coordinates(self, f):
        return (np.asarray(f) @ self.whitening)[self._rows, self._cols]

    def original_tangent(self, w):
        return self.rotation @ w @ self.rotation.conj().T

    def coordinate_tangent(self, w):
        return self.rotation.conj().T @ w @ self.rotation


# --------------------------------------------------------------------------------------------
# Davidson (src/davidson.jl)
# --------------------------------------------------------------------------------------------
@dataclass
class DavidsonOptions:
    tol: float = 1e-7
    maxops: int = 200
    maxspace: int = 8
    retain: int | None = None
    refresh: bool = True
    hermitian_tol: float = 1e-7
    stall_window: int = 20
    stall_minops: int = 40
    stall_rtol: float = 0.1
    stall_residual_limit: float | None = None
    stall_energy_atol: float = 1e-9

    def stall_limit(self):
        return self.stall_residual_limit if self.stall_residual_limit is not None else max(1e-5, 10 * self.tol)


def check_davidson(o: DavidsonOptions, nev):
    retain = o.retain if o.retain is not None else nev
    if nev < 1:
        raise ValueError("nev must be positive")
    if not (math.isfinite(o.tol) and o.tol > 0 and o.maxops >= 1 and math.isfinite(o.hermitian_tol)
            and o.hermitian_tol > 0):
        raise ValueError("invalid Davidson controls")
    if not (nev <= retain < o.maxspace):
        raise ValueError(f"require nev <= retain < maxspace (nev={nev}, retain={retain}, maxspace={o.maxspace})")
    if o.maxops < nev:
        raise ValueError("maxops must be at least nev")
    if o.stall_window < 0:
        raise ValueError("stall_window must be nonnegative")
    if o.stall_window > 0:
        limit = o.stall_limit()
        if not (o.stall_minops >= 1 and 0 < o.stall_rtol < 1 and o.stall_energy_atol >= 0):
            raise ValueError("invalid Davidson stagnation controls")
        if not (math.isfinite(limit) and limit >= o.tol):
            raise ValueError("stall_residual_limit must be at least tol")
    return retain


def _stalled(history, o: DavidsonOptions):
    w = o.stall_window
    if w <= 0 or not history:
        return False
    nev = len(history[-1]["energies"])
    full = [r for r in history if len(r["energies"]) == nev]
    n = len(full)
    if n < max(o.stall_minops, w + 1):
        return False
    if full[-1]["residual"] > o.stall_limit():
        return False
    old_best = min(r["residual"] for r in full[:n - w])
    best = min(r["residual"] for r in full)
    if best < (1 - o.stall_rtol) * old_best:
        return False
    recent = full[n - w - 1:]
    for k in range(nev):
        vals = [r["energies"][k] for r in recent]
        if max(vals) - min(vals) > o.stall_energy_atol:
            return False
    return True


def _orthogonalize(x, basis):
    for _ in range(2):
        for b in basis:
            x -= b * np.vdot(b, x)
    return x


def davidson(op, x0, *, nev=1, precondition=None, options: DavidsonOptions | None = None, callback=None,
             batch_op=None):
    """Lowest nev eigenpairs of a Hermitian map known through op(x) (src/davidson.jl:106-206):
    restarted Rayleigh-Ritz with fresh or thick restarts, exact residuals from stored actions, the
    stall rule and breakdown detection.  batch_op(list of x) -> list of actions, when given, applies
    the operator to the start vectors together."""
    gen = davidson_steps(x0, nev=nev, precondition=precondition, options=options, callback=callback,
                         blocks=batch_op is not None)
    apply = batch_op or (lambda xs: [op(x) for x in xs])
    try:
        request = next(gen)
        while True:
            request = gen.send(apply(request))
    except StopIteration as stop:
        return stop.value


def davidson_steps(x0, *, nev=1, precondition=None, options: DavidsonOptions | None = None, callback=None,
                   blocks=False):
    """The Davidson iteration of `davidson` as a generator: it yields lists of vectors to apply and
    receives their actions (so that several searches can share device calls); its return value is
    davidson's result.  With blocks, the start vectors of a (re)started space are applied together."""
    o = options or DavidsonOptions()
    retain = check_davidson(o, nev)
    precondition = precondition or (lambda r: r)
    starts = [x0] if np.ndim(x0) == 1 else list(x0)
    if not starts:
        raise ValueError("at least one start vector is required")
    pending = []
    for x in starts:
        y = np.array(x, dtype=complex)
        if not np.isfinite(y).all():
            raise ValueError("start vectors must be finite")
        _orthogonalize(y, pending)
        scale = np.linalg.norm(x)
        if scale > 0 and np.linalg.norm(y) > math.sqrt(np.finfo(float).eps) * scale:
            pending.append(y / np.linalg.norm(y))
    if not pending:
        raise ValueError("start vectors must be nonzero and independent")
    if len(pending) >= o.maxspace:
        raise ValueError(f"at most maxspace-1={o.maxspace - 1} start vectors are accepted")
    n = len(pending[0])
    if nev > n:
        raise ValueError(f"nev={nev} exceeds the dimension {n}")
    stale = [None] * len(pending)
    vs, avs, history = [], [], []
    cycle = 1
    prefetched = {}
    count = 0
    while count < o.maxops:
        count += 1
        x = pending.pop(0)
        previous = stale.pop(0)
        _orthogonalize(x, vs)
        x /= np.linalg.norm(x)
        if blocks and pending and not prefetched and len(vs) == 0:
            # the start block: orthonormalise all pending vectors now and apply them together
            block = [x]
            for y in pending:
                y = _orthogonalize(y.copy(), vs + block)
                block.append(y / np.linalg.norm(y))
            pending = block[1:]
            acts = yield block
            for y, a in zip(block[1:], acts[1:]):
                prefetched[id(y)] = a
            ax = np.asarray(acts[0], dtype=complex)
        elif id(x) in prefetched:
            ax = np.asarray(prefetched.pop(id(x)), dtype=complex)
        else:
            ax = np.asarray((yield [x])[0], dtype=complex)
        if len(ax) != n or not np.isfinite(ax).all():
            raise SolverFailure("invalid operator action")
        change = 0.0 if previous is None else float(np.linalg.norm(ax - previous))
        vs.append(x)
        avs.append(ax)
        V, AV = np.stack(vs, 1), np.stack(avs, 1)
        P = V.conj().T @ AV
        defect = np.linalg.norm(P - P.conj().T) / max(1.0, np.linalg.norm(P))
        if defect > o.hermitian_tol:
            raise SolverFailure(f"Davidson projected operator is not Hermitian: {defect}")
        evals, evecs = np.linalg.eigh((P + P.conj().T) / 2)
        m = min(nev, len(vs))
        U, AU = V @ evecs[:, :m], AV @ evecs[:, :m]
        for k in range(m):
            sc = np.linalg.norm(U[:, k])
            U[:, k] /= sc
            AU[:, k] /= sc
        energies = [float(np.vdot(U[:, k], AU[:, k]).real) for k in range(m)]
        Rr = AU - U * np.array(energies)
        residuals = [float(np.linalg.norm(Rr[:, k])) for k in range(m)]
        row = dict(count=count, cycle=cycle, dimension=len(vs), energies=energies, residuals=residuals,
                   energy=energies[0], residual=max(residuals), hermiticity=float(defect),
                   orthogonality=float(np.linalg.norm(V.conj().T @ V - np.eye(len(vs)))), refresh_change=change)
        history.append(row)
        if callback is not None:
            callback(row, U, AU)
        converged = m == nev and row["residual"] <= o.tol
        stalled = not converged and m == nev and _stalled(history, o)

        def finish(reason):
            return dict(energies=energies, vectors=[U[:, k].copy() for k in range(m)],
                        actions=[AU[:, k].copy() for k in range(m)], residuals=residuals,
                        converged=reason == "converged", stop_reason=reason, numops=count, history=history)
        if converged:
            return finish("converged")
        if stalled:
            return finish("stalled")
        if count == o.maxops:
            return finish("maxops")
        if pending:
            continue
        if len(vs) >= o.maxspace:
            keep = evecs[:, :min(retain, len(vs))]
            kept, keptact = V @ keep, AV @ keep
            cycle += 1
            if o.refresh and o.maxops - count < nev:
                return finish("maxops")
            if o.refresh:
                vs, avs = [], []
                pending = [kept[:, k].copy() for k in range(kept.shape[1])]
                stale = [keptact[:, k].copy() for k in range(kept.shape[1])]
                continue
            vs = [kept[:, k].copy() for k in range(kept.shape[1])]
            avs = [keptact[:, k].copy() for k in range(kept.shape[1])]

        def expansion(k):
            y = -np.asarray(precondition(Rr[:, k]), dtype=complex)
            if len(y) != n or not np.isfinite(y).all():
                raise SolverFailure("invalid preconditioned residual")
            _orthogonalize(y, vs)
            if np.linalg.norm(y) < 1e-13 * max(1.0, residuals[k]):
                y = Rr[:, k].copy()
                _orthogonalize(y, vs)
            return y if np.linalg.norm(y) > np.finfo(float).eps * max(1.0, residuals[k]) else None
        candidates = [k for k in range(m) if residuals[k] > o.tol] or [m - 1]
        y = None
        for k in candidates:
            y = expansion(k)
            if y is not None:
                break
        if y is None:
            if m == nev:
                return finish("breakdown")
            rng = np.random.default_rng(count)
            y = rng.standard_normal(n) + 1j * rng.standard_normal(n)
            _orthogonalize(y, vs)
        pending.append(y / np.linalg.norm(y))
        stale.append(None)
    raise AssertionError("unreachable")


# --------------------------------------------------------------------------------------------
# operator, preconditioner, start vectors, solve (src/excitation_eigsolve.jl:132-430)
# --------------------------------------------------------------------------------------------
class ExcitationOperator:
    """x -> [F(W(x)) B]_sector in white coordinates; fixed kinetic_nodes make it exactly linear.
    `calls` counts applications."""

    def __init__(self, ctx: ExcitationContext, coordinates: TangentCoordinates, model: ExcitationModel, *,
                 kinetic_nodes=256, kinetic_options=None):
        if ctx.state.D != coordinates.state.D:
            raise ValueError("context and coordinates differ")
        self.context, self.coordinates, self.model = ctx, coordinates, model
        self.kinetic_nodes = None if kinetic_nodes is None else check_kinetic_nodes(kinetic_nodes)
        self.kinetic_options = kinetic_options or KineticOptions()
        self.calls = 0

    def apply_many(self, xs):
        c = self.coordinates
        ws = [c.tangent_matrix(x) for x in xs]
        fs = excitation_hamiltonian_action(self.context, ws, model=self.model, kinetic_nodes=self.kinetic_nodes,
                                           kinetic_options=self.kinetic_options)
        self.calls += len(xs)
        return [c.covector_coordinates(f) for f in fs]

    def __call__(self, x):
        return self.apply_many([x])[0]

    def with_context(self, ctx, kinetic_nodes="same"):
        return ExcitationOperator(ctx, self.coordinates, self.model,
                                  kinetic_nodes=self.kinetic_nodes if kinetic_nodes == "same" else kinetic_nodes,
                                  kinetic_options=self.kinetic_options)


def kinetic_preconditioners(ops, *, nodes=8, floor=None, chunk=64, progress=None):
    """kinetic_preconditioner of several operators on one coordinate set (e.g. one per momentum):
    the coarse kinetic actions of every (operator, coordinate) pair share device calls of `chunk`
    tangents."""
    c = ops[0].coordinates
    n = len(c)
    model = ops[0].model
    if model.kinetic == (0.0, 0.0):
        raise ValueError("the kinetic preconditioner requires kinetic terms")
    nodes = check_kinetic_nodes(nodes)
    eye = np.eye(n, dtype=complex)
    jobs = [(k, j) for k in range(len(ops)) for j in range(n)]
    cols = [[None] * n for _ in ops]
    for j0 in range(0, len(jobs), chunk):
        part = jobs[j0:j0 + chunk]
        fs = action_batch([ops[k].context for k, _ in part], [c.tangent_matrix(eye[j]) for _, j in part], model, nodes,
                          _lib.EXC_KINETIC, krylov_options=PRECONDITIONER_KRYLOV)
        for (k, j), f in zip(part, fs):
            cols[k][j] = c.covector_coordinates(f)
        if progress is not None:
            progress(f"kinetic preconditioner: {min(j0 + chunk, len(jobs))}/{len(jobs)} coarse actions")
    lower = max(model.kinetic) if floor is None else float(floor)
    if not (math.isfinite(lower) and lower > 0):
        raise ValueError("preconditioner floor must be positive")
    out = []
    for k in range(len(ops)):
        m = np.stack(cols[k], 1)
        if not np.isfinite(m).all():
            raise SolverFailure("nonfinite kinetic preconditioner")
        vals, vecs = np.linalg.eigh((m + m.conj().T) / 2)
        vc, den = vecs.conj(), np.maximum(lower, vals)   # made once, not per application

        def apply(r, vecs=vecs, vc=vc, den=den):
            return vecs @ ((vc.T @ r) / den)
        out.append(dict(apply=apply, matrix=m, values=vals, vectors=vecs,
                        skew=float(np.linalg.norm(m - m.conj().T) / max(1.0, np.linalg.norm(m))), floor=lower,
                        nodes=nodes))
    return out


def kinetic_preconditioner(op: ExcitationOperator, *, nodes=8, floor=None, chunk=64):
    """r -> sum_k v_k <v_k, r> / max(floor, kappa_k) from the coarse kinetic matrix on `nodes`
    theta nodes (src/excitation_eigsolve.jl:181-201); one batched kinetic action per chunk."""
    return kinetic_preconditioners([op], nodes=nodes, floor=floor, chunk=chunk)[0]


def field_seed(op: ExcitationOperator):
    """White coordinates of the field insertion phi(0)|Psi>: W = 1 + [R, X] with
    (L^+ + ip) X = R + R^+ trace-projected (src/excitation_eigsolve.jl:203-210)."""
    s = op.context.state
    x = -action_dual_poisson(op.context, s.R + s.R.conj().T, shifted=True)
    return op.coordinates.tangent_coordinates(np.eye(s.D) + s.R @ x - x @ s.R)


def initial_vectors(op, start, nev, pre):
    c = op.coordinates
    n = len(c)

    def kinetic(k):
        return [] if pre is None else [pre["vectors"][:, j].copy() for j in range(min(k, n))]
    if isinstance(start, str):
        if start not in ("auto", "field", "kinetic", "random"):
            raise ValueError("start must be 'auto', 'field', 'kinetic', 'random', coordinates or tangents")
        if start in ("auto", "field"):
            x = field_seed(op)
            if np.linalg.norm(x) > 1e-6:
                return [x] + kinetic(nev - 1), "field"
            if start == "field":
                raise ValueError(f"the field insertion vanishes in the {c.sector} sector")
        if start in ("auto", "kinetic") and pre is not None:
            return kinetic(nev), "kinetic"
        if start == "kinetic":
            raise ValueError("start='kinetic' requires preconditioner='kinetic'")
        rng = np.random.default_rng(1234)
        return [rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(nev)], "random"
    starts = [start] if (isinstance(start, np.ndarray) and start.ndim in (1, 2)) else list(start)
    out = []
    for x in starts:
        x = np.asarray(x, dtype=complex)
        if x.ndim == 2:
            if x.shape != (c.state.D, c.state.D):
                raise ValueError("start tangent has the wrong size")
            out.append(c.tangent_coordinates(c.coordinate_tangent(x)))
        else:
            if len(x) != n:
                raise ValueError("start coordinates have the wrong length")
            out.append(x)
    return out, "supplied"


def refined_integration(p: IntegratorParameters) -> IntegratorParameters:
    """Window + 10, twice n_quad, tolerances / 10 (src/excitation_eigsolve.jl:250-252)."""
    return IntegratorParameters(n_quad=2 * p.n_quad, max_point=p.max_point + 10, atol=p.atol / 10, rtol=p.rtol / 10,
                                maxiters=p.maxiters)


def fresh_level(op, x):
    h = op(x)
    value = np.vdot(x, h) / np.vdot(x, x)
    return dict(energy=float(value.real), imaginary=float(value.imag),
                residual=float(np.linalg.norm(h - value.real * x) / np.linalg.norm(x)), action=h)


def _apply_many(ops, requests):
    """Apply the operators to requested vectors in one device call: requests = [(operator index,
    vector)]; returns the actions in order."""
    if not requests:
        return []
    c = ops[0].coordinates
    op0 = ops[0]
    ws = [c.tangent_matrix(x) for _, x in requests]
    fs = excitation_hamiltonian_action([ops[k].context for k, _ in requests], ws, model=op0.model,
                                       kinetic_nodes=op0.kinetic_nodes, kinetic_options=op0.kinetic_options)
    for k, _ in requests:
        ops[k].calls += 1
    return [c.covector_coordinates(f) for f in fs]


def _check_kinetic_options(o: KineticOptions):
    """The quadrature controls check of src/excitations.jl:89-91."""
    if not (o.minnodes >= 2 and o.maxnodes >= 2 * o.minnodes and np.isfinite(o.atol) and np.isfinite(o.rtol)
            and o.atol > 0 and o.rtol >= 0):
        raise ValueError("invalid kinetic quadrature controls")


def _pair_elements(ctxs, us, ws, model, kinetic_nodes, kinetic_options):
    """Connected elements <Phi_p(u)|H - E0|Phi_p(w)> of pairs (one context each), the kinetic grid fixed
    or refined by node doubling until every pair's kinetic sum has converged
    (excitation_element / excitation_kinetic, src/excitations.jl:88-119)."""
    if kinetic_nodes is not None:
        return excitation_element(ctxs, us, ws, model=model, kinetic_nodes=kinetic_nodes)
    o = kinetic_options or KineticOptions()
    _check_kinetic_options(o)
    pot = excitation_element(ctxs, us, ws, model=model, kinetic_nodes=2, parts=_lib.EXC_POTENTIAL)
    n = o.minnodes
    prev = excitation_element(ctxs, us, ws, model=model, kinetic_nodes=n, parts=_lib.EXC_KINETIC)
    while 2 * n <= o.maxnodes:
        n *= 2
        cur = excitation_element(ctxs, us, ws, model=model, kinetic_nodes=n, parts=_lib.EXC_KINETIC)
        if not np.isfinite(cur).all():
            raise SolverFailure("nonfinite excitation kinetic quadrature")
        if np.all(np.abs(cur - prev) <= o.atol + o.rtol * np.abs(cur)):
            return cur + pot
        prev = cur
    raise SolverFailure(f"excitation kinetic quadrature did not converge by {n} nodes")


def _elements(ctxs, tangents, model, kinetic_nodes, kinetic_options):
    """Diagonal connected elements <W|H - E0|W> of tangents (one context each)."""
    return _pair_elements(ctxs, tangents, tangents, model, kinetic_nodes, kinetic_options)


def excitation_matrix_element(s: State, p: float, u, w, *, model: ExcitationModel | None = None,
                              environment_tol: float = 1e-10, kinetic_nodes=None,
                              kinetic_options: KineticOptions | None = None):
    """excitation_matrix_element(s, p, u, w) (src/excitations.jl:122-130): the connected element
    <Phi_p(u)|H - E0|Phi_p(w)> per 2 pi delta(0) of horizontal tangents u, w (V = -R^+ W), a complex
    sesquilinear form.  kinetic_nodes=None refines the theta grid by node doubling; an integer fixes
    it.  `u` and `w` may be equal-length lists of tangents: one batched device call per grid, and an
    array of elements is returned."""
    many = isinstance(w, (list, tuple))
    us, ws = (list(u), list(w)) if many else ([u], [w])
    if kinetic_nodes is not None:
        check_kinetic_nodes(kinetic_nodes)
    ctx = ExcitationContext(s, p, tol=environment_tol)
    try:
        vals = _pair_elements([ctx] * len(ws), us, ws, model or ExcitationModel.of(s), kinetic_nodes,
                              kinetic_options)
    finally:
        ctx.close()
    return vals if many else complex(vals[0])


def solve_levels_many(coordinates, ctxs, model, *, nev=1, kinetic_nodes=256, kinetic_options=None,
                      preconditioner="kinetic", preconditioner_nodes=8, start="auto", davidson_options=None,
                      validate=True, refined=None, refined_kinetic_nodes=None, validation_tol=1e-7,
                      acceptance_residual=1e-5, callback=None, context_seconds=0.0, progress=None,
                      preconditioner_chunk=64):
    """The iterative lowest-level solve (src/excitation_eigsolve.jl:302-401, without refinement
    steps) for one context per momentum on shared coordinates, the Davidson searches in lockstep:
    every round applies the pending vectors of all searches in one device call.  Validation computes
    the four energies of every level (fresh actions and bilinear elements, standard and refined
    settings; each kind in one device call).  Returns one result per context.
    progress(message), when given, is called after every device call."""
    n = len(coordinates)
    if not (isinstance(nev, (int, np.integer)) and 1 <= nev <= n):
        raise ValueError(f"nev={nev} must lie between 1 and the sector dimension {n}")
    options = davidson_options or DavidsonOptions(maxspace=max(8, 2 * nev + 2))
    check_davidson(options, nev)
    if not (preconditioner in ("kinetic", "none") or callable(preconditioner)):
        raise ValueError("preconditioner must be 'kinetic', 'none', or a function")
    if not (acceptance_residual > 0 and validation_tol > 0):
        raise ValueError("acceptance_residual and validation_tol must be positive")
    ops = [ExcitationOperator(ctx, coordinates, model, kinetic_nodes=kinetic_nodes, kinetic_options=kinetic_options)
           for ctx in ctxs]
    tick = time.time()
    pres = (kinetic_preconditioners(ops, nodes=preconditioner_nodes, chunk=preconditioner_chunk, progress=progress)
            if preconditioner == "kinetic" else [None] * len(ops))
    applies = [p["apply"] if p is not None else (None if preconditioner == "none" else preconditioner) for p in pres]
    pre_seconds = time.time() - tick
    starts = [initial_vectors(op, start, nev, pre) for op, pre in zip(ops, pres)]
    tick = time.time()
    gens = [davidson_steps(sv, nev=nev, precondition=ap, options=options, callback=callback, blocks=True)
            for (sv, _), ap in zip(starts, applies)]
    results = [None] * len(ops)
    pending = {}
    for k, gen in enumerate(gens):
        try:
            pending[k] = next(gen)
        except StopIteration as stop:
            results[k] = stop.value
    rounds = 0
    while pending:
        flat = [(k, x) for k in sorted(pending) for x in pending[k]]
        acts = _apply_many(ops, flat)
        rounds += 1
        if progress is not None:
            progress(f"Davidson round {rounds}: {len(flat)} actions for {len(pending)} searches")
        pos = 0
        for k in sorted(pending):
            cnt = len(pending[k])
            mine, pos = acts[pos:pos + cnt], pos + cnt
            try:
                pending[k] = gens[k].send(mine)
            except StopIteration as stop:
                results[k] = stop.value
                del pending[k]
    solve_seconds = time.time() - tick
    tick = time.time()
    validation = [None] * len(ops)
    refined_ctxs = [None] * len(ops)
    if validate:
        rp = refined or refined_integration(ctxs[0].integrator)
        rn = refined_kinetic_nodes if refined_kinetic_nodes is not None else (
            None if ops[0].kinetic_nodes is None else 2 * ops[0].kinetic_nodes)
        refined_ctxs = [ctx.with_settings(integrator=rp) for ctx in ctxs]
        refs = [op.with_context(rc, kinetic_nodes=rn) for op, rc in zip(ops, refined_ctxs)]
        flat = [(k, v) for k in range(len(ops)) for v in results[k]["vectors"]]
        std = _apply_many(ops, flat)
        fine = _apply_many(refs, flat)
        # the bilinear energies: connected elements <W|H - E0|W> / tr(W rho W^+) at both settings
        tang = [coordinates.tangent_matrix(x) for _, x in flat]
        bil = _elements([ops[k].context for k, _ in flat], tang, model, ops[0].kinetic_nodes,
                        ops[0].kinetic_options)
        rbil = _elements([refs[k].context for k, _ in flat], tang, model, refs[0].kinetic_nodes,
                         refs[0].kinetic_options)
        for k in range(len(ops)):
            validation[k] = []
        for (k, x), hs, hf, wt, eb, erb in zip(flat, std, fine, tang, bil, rbil):
            levels = []
            for h in (hs, hf):
                value = np.vdot(x, h) / np.vdot(x, x)
                levels.append((float(value.real), float(value.imag),
                               float(np.linalg.norm(h - value.real * x) / np.linalg.norm(x))))
            (es, ims, rs), (ef, imf, rf) = levels
            norm = tangent_metric(coordinates.state, wt, wt)
            vb, vrb = eb / norm, erb / norm
            values = (es, ef, vb.real, vrb.real)
            spread = max(values) - min(values)
            imag = max(abs(ims), abs(imf), abs(vb.imag), abs(vrb.imag))
            residual = max(rs, rf)
            validation[k].append(dict(standard_energy=es, standard_residual=rs, refined_energy=ef, refined_residual=rf,
                                      bilinear_energy=float(vb.real), refined_bilinear_energy=float(vrb.real),
                                      energy_spread=spread, max_imaginary=imag,
                                      residual_target_met=residual <= validation_tol,
                                      consistent=spread <= validation_tol and imag <= validation_tol,
                                      accepted=spread <= validation_tol and imag <= validation_tol
                                      and residual <= acceptance_residual))
    validation_seconds = time.time() - tick
    out = []
    for k, (op, ctx, res, pre) in enumerate(zip(ops, ctxs, results, pres)):
        coefficients = [v.copy() for v in res["vectors"]]
        val = validation[k] if validate else None
        accepted = len(coefficients) == nev and (res["converged"] if val is None else all(v["accepted"] for v in val))
        out.append(dict(
            energies=list(res["energies"]), coefficients=coefficients,
            vectors=[coordinates.original_tangent(coordinates.tangent_matrix(x)) for x in coefficients],
            residuals=list(res["residuals"]), converged=res["converged"], stop_reason=res["stop_reason"],
            accepted=accepted, nev=nev, validation_tol=validation_tol, acceptance_residual=acceptance_residual,
            numops=res["numops"], action_calls=op.calls, history=res["history"], momentum=ctx.momentum,
            sector=coordinates.sector, parity=coordinates.parity, gauge=coordinates.gauge, dimension=n,
            metric_rank=coordinates.metric_rank, start=starts[k][1], davidson=options,
            preconditioner=None if pre is None else dict(nodes=pre["nodes"], skew=pre["skew"], floor=pre["floor"],
                                                          minimum=float(pre["values"].min())),
            validation=val,
            validated_energies=None if val is None else [v["refined_bilinear_energy"] for v in val],
            integration=ctx.integrator,
            refined_integration=None if refined_ctxs[k] is None else refined_ctxs[k].integrator,
            kinetic_nodes=op.kinetic_nodes, device_rounds=rounds,
            timings=dict(context=context_seconds, preconditioner=pre_seconds, solve=solve_seconds,
                         validation=validation_seconds),
            model=model.name, coordinates=coordinates))
    return out


def solve_levels(coordinates, ctx, model, **kw):
    """The iterative lowest-level solve at one momentum (src/excitation_eigsolve.jl:302-401); the
    refinement steps (refine_steps) of the reference are not built."""
    return solve_levels_many(coordinates, [ctx], model, **kw)[0]


def excitation_eigsolve(s: State, p: float = 0.0, *, model: ExcitationModel | None = None, sector="auto",
                        parity="auto", parity_tol=1e-12, gauge="density", metric_tol=0.0, environment_tol=1e-10,
                        max_bond=64, **kw):
    """Lowest nev tangent-space excitation energies at momentum p by preconditioned Davidson on
    device Hamiltonian actions (src/excitation_eigsolve.jl:430-447)."""
    if not math.isfinite(p):
        raise ValueError("invalid momentum")
    if s.D > max_bond:
        raise ValueError(f"excitation_eigsolve is limited to D<={max_bond}")
    tick = time.time()
    model = model or ExcitationModel.of(s)
    c = TangentCoordinates(s, sector=sector, parity=parity, parity_tol=parity_tol, gauge=gauge, metric_tol=metric_tol)
    ctx = ExcitationContext(c.state, p, tol=environment_tol)
    return solve_levels(c, ctx, model, context_seconds=time.time() - tick, **kw)


def iterative_dispersion(s: State, momenta, *, model=None, sector="auto", parity="auto", parity_tol=1e-12,
                         gauge="density", metric_tol=0.0, environment_tol=1e-10, max_bond=64, start="auto",
                         strict=True, lockstep=False, **kw):
    """excitation_eigsolve at every momentum on shared coordinates (src/excitation_eigsolve.jl:453-476).
    Sequentially, each momentum warm-starts from the previous eigenvectors, as the reference does;
    with lockstep=True every momentum starts from `start` and the Davidson searches of all momenta
    share their device calls (one action per momentum per call)."""
    if s.D > max_bond:
        raise ValueError(f"excitation_eigsolve is limited to D<={max_bond}")
    model = model or ExcitationModel.of(s)
    c = TangentCoordinates(s, sector=sector, parity=parity, parity_tol=parity_tol, gauge=gauge, metric_tol=metric_tol)
    if lockstep:
        ctxs = [ExcitationContext(c.state, float(p), tol=environment_tol) for p in momenta]
        results = solve_levels_many(c, ctxs, model, start=start, **kw)
        if strict:
            for r in results:
                accepted_energies(r)
        return results
    results = []
    for k, p in enumerate(momenta):
        ctx = ExcitationContext(c.state, float(p), tol=environment_tol)
        seed = start if (k == 0 or not (isinstance(start, str) and start == "auto")) else results[-1]["coefficients"]
        results.append(solve_levels(c, ctx, model, start=seed, **kw))
        if strict:
            accepted_energies(results[-1])
    return results


def reported_energies(r):
    return r["energies"] if r["validated_energies"] is None else r["validated_energies"]


def accepted_energies(r):
    if r["accepted"]:
        return reported_energies(r)
    raise SolverFailure(f"iterative excitation levels at p={r['momentum']} not accepted (stop: {r['stop_reason']})")


def dispersion_relation(s: State, momenta, *, nev=1, strict=True, **kw):
    """Energy-sorted levels at each momentum by the iterative solver (src/excitations.jl:227-246)."""
    ps = [float(p) for p in momenta]
    if not ps or not all(math.isfinite(p) for p in ps):
        raise ValueError("momenta must be nonempty and finite")
    results = iterative_dispersion(s, ps, nev=nev, strict=strict, **kw)
    levels = [accepted_energies(r) if strict else reported_energies(r) for r in results]
    if not all(len(l) == nev for l in levels):
        raise SolverFailure(f"some momenta returned fewer than nev={nev} levels")
    return dict(momenta=ps, energies=np.array(levels), spectra=results, accepted=[r["accepted"] for r in results])


def relativistic_diagnostics(dispersion, *, branch=1, target_speed=1.0):
    """E(p)^2 - c^2 p^2 and the least-squares fit E^2 = intercept + c^2 p^2 (src/excitations.jl:250-277)."""
    if not (math.isfinite(target_speed) and target_speed > 0):
        raise ValueError("target_speed must be finite and positive")
    e = np.asarray(dispersion["energies"])
    if not 1 <= branch <= e.shape[1]:
        raise ValueError("invalid branch")
    ps = np.asarray(dispersion["momenta"], dtype=float)
    if len(np.unique(ps ** 2)) < 2:
        raise ValueError("at least two distinct squared momenta are required")
    energies = e[:, branch - 1]
    if (energies < 0).any():
        raise ValueError("negative excitation energies cannot be interpreted as a relativistic branch")
    p2, y = ps ** 2, energies ** 2
    a = np.stack([np.ones_like(ps), p2], 1)
    (intercept, c2), *_ = np.linalg.lstsq(a, y, rcond=None)
    invariant = y - target_speed ** 2 * p2
    return dict(invariant_mass_squared=invariant, intercept=float(intercept), c_squared=float(c2),
                fit_residual=float(np.linalg.norm(y - a @ np.array([intercept, c2])) / max(np.linalg.norm(y), 1e-300)),
                invariant_spread=float(invariant.max() - invariant.min()))
