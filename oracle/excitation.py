"""CPU restatement of RcmpsCore's tangent-space excitation subsystem (test infrastructure).

Only tests/ import this module, and only as the checker of the GPU excitation action.  It
follows the reference's dense small-D route function by function:

- ExcitationContext: dense transfer matrix, complex Schur form, mixing check, pinned LU factors
  of -L, -L + ip, -L - ip (src/excitation_environments.jl:18-49);
- Poisson solves and the tangent pair environments (:75-136);
- the connected matrix element: the four-block transfer ODE and the kinetic resolvent sum
  (src/excitations.jl:25-135);
- the reverse-contracted Hamiltonian action: forward/adjoint action ODEs with their pullback, the
  kinetic theta sum with its boundary covectors (src/excitation_action.jl:1-237);
- white tangent coordinates of one parity sector (src/excitation_eigsolve.jl:1-130).

Matrices are complex128, row-major.  The transfer matrix acts on row-major flattened matrices:
T = kron(Q, I) + kron(I, conj Q) + kron(R, conj R), so T vec(X) = vec(L(X)) and T^H is the
matrix of the adjoint map L*.  ODEs run with SciPy's DOP853 (published Dormand-Prince 8(5,3)
with its dense output at the quadrature nodes), stopping at t = 0 as the reference does.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla
from scipy.integrate import solve_ivp

from . import quadrature as _quadrature, source as _source


class ExcitationFailure(RuntimeError):
    """SolverFailure of the reference."""


# --------------------------------------------------------------------------------------------
# observables (src/observables.jl:7-29, 255-256)
# --------------------------------------------------------------------------------------------
@dataclass(frozen=True)
class Edge:
    src: int          # 0-based channel indices
    dst: int
    left: bool
    coefficient: float


@dataclass(frozen=True)
class Spec:
    channels: int
    edges: tuple
    commutator: bool
    kinetic: bool


def monomial_spec(n):
    edges = []
    for k in range(1, n + 1):
        edges += [Edge(k - 1, k, True, float(k)), Edge(k - 1, k, False, float(k))]
    return Spec(n + 1, tuple(edges), False, False)


def kinetic_spec(commutator):
    s2 = math.sqrt(2)
    return Spec(4, (Edge(0, 1, True, s2), Edge(0, 2, False, s2), Edge(2, 3, True, s2), Edge(1, 3, False, s2)),
                bool(commutator), True)


def phi4_terms(g):
    """ExcitationModel(:phi4, g) (src/excitations.jl:10-18): zero weights are dropped."""
    terms = [(1.0, kinetic_spec(False)), (1.0, kinetic_spec(True)), (float(g), monomial_spec(4))]
    return [t for t in terms if t[0] != 0]


def transformed_source(t):
    return _source(t)


def kinetic_grid(n):
    """src/kinetic.jl:18-22: Gauss-Legendre on (0, pi/2), weights including the factor 2/pi."""
    x, w = np.polynomial.legendre.leggauss(int(n))
    return np.pi / 4 * (x + 1), w / 2


# --------------------------------------------------------------------------------------------
# the state and its context
# --------------------------------------------------------------------------------------------
@dataclass
class XState:
    R: np.ndarray
    Q: np.ndarray
    rho: np.ndarray
    n_quad: int = 128
    max_point: float = 20.0
    atol: float = 1e-11
    rtol: float = 1e-10

    @property
    def D(self):
        return len(self.R)

    def settings(self, **kw):
        out = XState(self.R, self.Q, self.rho, self.n_quad, self.max_point, self.atol, self.rtol)
        for k, v in kw.items():
            setattr(out, k, v)
        return out


def lindblad(s, x):
    return s.Q @ x + x @ s.Q.conj().T + s.R @ x @ s.R.conj().T


def lindblad_adjoint(s, x):
    return s.Q.conj().T @ x + x @ s.Q + s.R.conj().T @ x @ s.R


def transfer_matrix(Q, R):
    d = len(Q)
    eye = np.eye(d)
    return np.kron(Q, eye) + np.kron(eye, Q.conj()) + np.kron(R, R.conj())


@dataclass
class Context:
    state: XState
    momentum: float
    T: np.ndarray
    S: np.ndarray       # upper-triangular Schur factor, T = Z S Z^H
    Z: np.ndarray
    poisson: tuple
    ket: tuple
    bra: tuple
    tol: float


def context(s: XState, p: float, tol=1e-10, max_bond=16) -> Context:
    """ExcitationContext(s, p) (src/excitation_environments.jl:18-40)."""
    if not (math.isfinite(p) and math.isfinite(tol) and tol > 0):
        raise ValueError("invalid momentum or environment tolerance")
    if s.D > max_bond:
        raise ValueError("dense excitation reference is limited by max_bond")
    T = transfer_matrix(s.Q, s.R)
    S, Z = sla.schur(T, output="complex")
    values = np.diag(S)
    scale = max(1.0, np.linalg.norm(T))
    small = np.abs(values) <= tol * scale
    if not (np.count_nonzero(small) == 1 and np.all(small | (values.real < -tol * scale))):
        raise ExcitationFailure("excitation environments require a unique, mixing transfer fixed point")
    d = s.D
    P = np.outer(s.rho.ravel(), np.eye(d).ravel())
    n = d * d
    poisson = sla.lu_factor(-T + P)
    if p == 0:
        ket = bra = poisson
    else:
        ket = sla.lu_factor(-T + 1j * p * np.eye(n) + P)
        bra = sla.lu_factor(-T - 1j * p * np.eye(n) + P)
    return Context(s, float(p), T, S, Z, poisson, ket, bra, float(tol))


def with_settings(ctx: Context, s: XState) -> Context:
    """The same factors with other integration settings (src/excitation_environments.jl:58-71)."""
    return Context(s, ctx.momentum, ctx.T, ctx.S, ctx.Z, ctx.poisson, ctx.ket, ctx.bra, ctx.tol)


def excitation_poisson(ctx, b, shift=0):
    """(-L + i shift p + |rho><I|) x = b - tr(b) rho, trace-projected (:75-85)."""
    s = ctx.state
    rhs = b - np.trace(b) * s.rho
    f = ctx.poisson if shift == 0 else ctx.ket if shift == 1 else ctx.bra
    x = sla.lu_solve(f, rhs.ravel()).reshape(s.D, s.D)
    x = x - np.trace(x) * s.rho
    residual = -lindblad(s, x) + (1j * shift * ctx.momentum) * x - rhs
    if not np.linalg.norm(residual) <= 20 * ctx.tol * max(1.0, np.linalg.norm(rhs)):
        raise ExcitationFailure("tangent boundary solve residual too large")
    return x


def action_dual_poisson(ctx, b, shifted=False):
    """(-L* - i p [shifted] + |I><rho|) x = b - tr(rho b) I, projected (src/excitation_action.jl:85-94)."""
    s = ctx.state
    eye = np.eye(s.D)
    rhs = b - np.trace(s.rho @ b) * eye
    f = ctx.ket if shifted else ctx.poisson
    x = sla.lu_solve(f, rhs.ravel(), trans=2).reshape(s.D, s.D)
    x = x - np.trace(s.rho @ x) * eye
    residual = -lindblad_adjoint(s, x) - (1j * ctx.momentum * x if shifted else 0) - rhs
    if not np.linalg.norm(residual) <= 20 * ctx.tol * max(1.0, np.linalg.norm(rhs)):
        raise ExcitationFailure("excitation action boundary residual too large")
    return x


def tangent_metric(s, u, w):
    return np.trace(w @ s.rho @ u.conj().T)


def ket_insertion(s, v, w, x):
    return v @ x + w @ x @ s.R.conj().T


def bra_insertion(s, v, u, x):
    return x @ v.conj().T + s.R @ x @ u.conj().T


# --------------------------------------------------------------------------------------------
# the pair environments and the connected matrix element
# --------------------------------------------------------------------------------------------
@dataclass
class Pair:
    ctx: Context
    w: np.ndarray
    u: np.ndarray
    v: np.ndarray
    vbra: np.ndarray
    overlap: complex
    right: np.ndarray      # (4, D, D): channels 00, ket, bra, ket+bra
    left00: np.ndarray


def pair(ctx, u, w) -> Pair:
    """ExcitationPair (src/excitation_environments.jl:105-127)."""
    s = ctx.state
    u = np.asarray(u, np.complex128)
    w = np.asarray(w, np.complex128)
    if u.shape != s.R.shape or w.shape != s.R.shape:
        raise ValueError("tangent dimensions do not match the state")
    Rh = s.R.conj().T
    v, vbra = -Rh @ w, -Rh @ u
    n = tangent_metric(s, u, w)
    right = np.zeros((4, s.D, s.D), np.complex128)
    right[0] = s.rho
    right[1] = excitation_poisson(ctx, ket_insertion(s, v, w, s.rho), 1)
    right[2] = excitation_poisson(ctx, bra_insertion(s, vbra, u, s.rho), -1)
    f = ket_insertion(s, v, w, right[2]) + bra_insertion(s, vbra, u, right[1]) + w @ s.rho @ u.conj().T
    right[3] = excitation_poisson(ctx, f)
    b = u.conj().T @ w - n * np.eye(s.D)
    h = sla.lu_solve(ctx.poisson, b.ravel(), trans=2).reshape(s.D, s.D)
    h = h - np.trace(s.rho @ h) * np.eye(s.D)
    if not np.linalg.norm(-lindblad_adjoint(s, h) - b) <= 20 * ctx.tol * max(1.0, np.linalg.norm(b)):
        raise ExcitationFailure("tangent right-tail solve residual too large")
    return Pair(ctx, w, u, v, vbra, n, right, h)


def excitation_transfer(x, pr: Pair):
    """The connected four-block transfer map (src/excitation_environments.jl:143-163)."""
    s, p = pr.ctx.state, pr.ctx.momentum
    Rh = s.R.conj().T
    out = np.stack([lindblad(s, x[k]) for k in range(4)])
    out[1] -= 1j * p * x[1]
    out[2] += 1j * p * x[2]
    for src, dst in ((0, 1), (2, 3)):
        out[dst] += pr.v @ x[src] + pr.w @ x[src] @ Rh
    for src, dst in ((0, 2), (1, 3)):
        out[dst] += x[src] @ pr.vbra.conj().T + s.R @ x[src] @ pr.u.conj().T
    out[3] += pr.w @ x[0] @ pr.u.conj().T - pr.overlap * x[0]
    return out


def excitation_contract(pr: Pair, x):
    return np.trace(x[3]) + np.trace(pr.left00 @ x[0])


def excitation_operators(pr: Pair, commutator):
    s, p = pr.ctx.state, pr.ctx.momentum
    if commutator:
        Q, R = s.Q, s.R
        c = Q @ R - R @ Q
        cw = Q @ pr.w - pr.w @ Q + pr.v @ R - R @ pr.v - 1j * p * pr.w
        cu = Q @ pr.u - pr.u @ Q + pr.vbra @ R - R @ pr.vbra - 1j * p * pr.u
        return c, cw, cu
    return s.R, pr.w, pr.u


def excitation_source_add(out, x, ops, left, a):
    c, cw, cu = ops
    if left:
        out += a * (c @ x)
        out[1] += a * (cw @ x[0])
        out[3] += a * (cw @ x[2])
    else:
        out += a * (x @ c.conj().T)
        out[2] += a * (x[0] @ cu.conj().T)
        out[3] += a * (x[1] @ cu.conj().T)
    return out


def _integrate(rhs, u0, s, t_eval=None):
    """Integrate du/dt = rhs(t, u) over (-a, a), stopping at 0; returns the states at the sorted
    t_eval points (DOP853 dense output) and the final state."""
    a = math.sqrt(s.max_point)
    shape = u0.shape
    y = u0.ravel().astype(np.complex128)
    f = lambda t, yy: rhs(t, yy.reshape(shape)).ravel()
    samples = []
    for lo, hi in ((-a, 0.0), (0.0, a)):
        te = [] if t_eval is None else [t for t in t_eval if (t < 0 if lo < 0 else t >= 0)]
        sol = solve_ivp(f, (lo, hi), y, method="DOP853", atol=s.atol, rtol=s.rtol, dense_output=bool(te))
        if sol.status != 0:
            raise ExcitationFailure(f"excitation ODE failed: {sol.message}")
        samples += [sol.sol(t).reshape(shape) for t in te]
        y = sol.y[:, -1]
    final = y.reshape(shape)
    if not np.all(np.isfinite(final)):
        raise ExcitationFailure("nonfinite excitation trajectory")
    return samples, final


def excitation_observable(pr: Pair, spec: Spec):
    """src/excitations.jl:38-53: connected element of one observable by the four-block ODE."""
    s = pr.ctx.state
    ops = excitation_operators(pr, spec.commutator)

    def rhs(t, u):
        jac, j = 2 * abs(t), transformed_source(t)
        du = np.stack([excitation_transfer(u[k], pr) for k in range(spec.channels)]) * jac
        for e in spec.edges:
            excitation_source_add(du[e.dst], u[e.src], ops, e.left, e.coefficient * j)
        return du

    u0 = np.zeros((spec.channels, 4, s.D, s.D), np.complex128)
    u0[0] = pr.right
    _, final = _integrate(rhs, u0, s)
    return excitation_contract(pr, final[-1])


def _resolvent_factor(ctx, z, shift):
    m = -z * ctx.S
    m[np.diag_indices_from(m)] += 1 + 1j * z * shift * ctx.momentum
    return m


def _solve_block(ctx, rhs, factor, adj=False):
    d = ctx.state.D
    y = ctx.Z.conj().T @ rhs.ravel()
    if adj:
        y = sla.solve_triangular(factor, y, lower=False, trans=2)
    else:
        y = sla.solve_triangular(factor, y, lower=False)
    return (ctx.Z @ y).reshape(d, d)


def excitation_resolvent(pr: Pair, b, z):
    """(src/excitation_environments.jl:203-218) (I - z M)^-1 b for the four-block transfer M."""
    ctx, s = pr.ctx, pr.ctx.state
    f0 = _resolvent_factor(ctx, z, 0)
    fk = _resolvent_factor(ctx, z, 1)
    fb = _resolvent_factor(ctx, z, -1)
    x = np.empty_like(b)
    x[0] = _solve_block(ctx, b[0], f0)
    x[1] = _solve_block(ctx, b[1] + z * ket_insertion(s, pr.v, pr.w, x[0]), fk)
    x[2] = _solve_block(ctx, b[2] + z * bra_insertion(s, pr.vbra, pr.u, x[0]), fb)
    rhs = b[3] + z * (ket_insertion(s, pr.v, pr.w, x[2]) + bra_insertion(s, pr.vbra, pr.u, x[1]) +
                      pr.w @ x[0] @ pr.u.conj().T - pr.overlap * x[0])
    x[3] = _solve_block(ctx, rhs, f0)
    return x


def excitation_kinetic_sum(pr: Pair, terms, n):
    """src/excitations.jl:55-77."""
    result = 0j
    sources = []
    for weight, spec in terms:
        ops = excitation_operators(pr, spec.commutator)
        bs = [excitation_source_add(np.zeros_like(pr.right), pr.right, ops, left, 1) for left in (False, True)]
        sources.append((weight, ops, bs))
    for theta, wn in zip(*kinetic_grid(n)):
        z = math.cos(theta)
        for weight, ops, bs in sources:
            for left, b in zip((False, True), bs):
                x = excitation_resolvent(pr, b, z)
                y = excitation_source_add(np.zeros_like(x), x, ops, not left, 1)
                result += weight * wn / 2 * excitation_contract(pr, y)
    return result


@dataclass
class KineticOptions:
    minnodes: int = 32
    maxnodes: int = 512
    atol: float = 1e-11
    rtol: float = 1e-10


def excitation_kinetic(pr, terms, o=KineticOptions()):
    """src/excitations.jl:82-95: node doubling until two levels agree."""
    n = o.minnodes
    previous = excitation_kinetic_sum(pr, terms, n)
    while 2 * n <= o.maxnodes:
        n *= 2
        current = excitation_kinetic_sum(pr, terms, n)
        if abs(current - previous) <= o.atol + o.rtol * abs(current):
            return current
        previous = current
    raise ExcitationFailure(f"excitation kinetic quadrature did not converge by {n} nodes")


def excitation_element(pr: Pair, terms, kinetic_backend="resolvent", kinetic_nodes=None,
                       kinetic_options=KineticOptions()):
    """src/excitations.jl:98-113."""
    result = 0j
    kinetics = [t for t in terms if t[1].kinetic]
    if kinetic_backend == "resolvent" and kinetics:
        result += (excitation_kinetic(pr, kinetics, kinetic_options) if kinetic_nodes is None
                   else excitation_kinetic_sum(pr, kinetics, kinetic_nodes))
    for weight, spec in terms:
        if kinetic_backend == "resolvent" and spec.kinetic:
            continue
        result += weight * excitation_observable(pr, spec)
    return result


# --------------------------------------------------------------------------------------------
# the reverse-contracted Hamiltonian action (src/excitation_action.jl)
# --------------------------------------------------------------------------------------------
def action_operators(ctx, w, v, commutator):
    s = ctx.state
    if not commutator:
        return s.R, w
    Q, R = s.Q, s.R
    return Q @ R - R @ Q, Q @ w - w @ Q + v @ R - R @ v - 1j * ctx.momentum * w


def bra_transfer_pull(s, a, x):
    return (a @ s.R - s.R @ a) @ x


def ket_trace_dual(s, v, w, a):
    return a @ v + s.R.conj().T @ a @ w


def bra_source_pull(ctx, x, commutator):
    s = ctx.state
    if not commutator:
        return x
    Qh, R, Rh = s.Q.conj().T, s.R, s.R.conj().T
    return Qh @ x - x @ Qh - R @ x @ Rh + R @ Rh @ x + 1j * ctx.momentum * x


def action_boundaries(ctx, w, v, rw, a0, a1, xout):
    """src/excitation_action.jl:95-101."""
    s = ctx.state
    d = action_dual_poisson(ctx, a1)
    e = action_dual_poisson(ctx, a0 + ket_trace_dual(s, v, w, d), shifted=True)
    return (bra_transfer_pull(s, e, s.rho) + bra_transfer_pull(s, d, rw) + d @ w @ s.rho +
            w @ excitation_poisson(ctx, xout))


def action_trajectory(ctx, w, v, rw, spec, points, adj):
    """src/excitation_action.jl:13-62: samples at `points` and the final state; u[k, 0/1] = x/y
    (forward) or a0/a1 (adjoint) of channel k."""
    s, p = ctx.state, ctx.momentum
    Rh = s.R.conj().T
    c, cw = action_operators(ctx, w, v, spec.commutator)
    ch = c.conj().T

    def rhs(t, u):
        jac, j = 2 * abs(t), transformed_source(t)
        du = np.empty_like(u)
        for k in range(spec.channels):
            x, y = u[k, 0], u[k, 1]
            if adj:
                du[k, 0] = lindblad_adjoint(s, x) + 1j * p * x + y @ v + Rh @ y @ w
                du[k, 1] = lindblad_adjoint(s, y)
            else:
                du[k, 0] = lindblad(s, x)
                du[k, 1] = lindblad(s, y) - 1j * p * y + v @ x + w @ x @ Rh
        du *= jac
        for e in spec.edges:
            a = e.coefficient * j
            src, dst = (e.dst, e.src) if adj else (e.src, e.dst)
            x, y = u[src, 0], u[src, 1]
            if adj and e.left:
                du[dst, 0] += a * (x @ c + y @ cw)
                du[dst, 1] += a * (y @ c)
            elif adj:
                du[dst, 0] += a * (ch @ x)
                du[dst, 1] += a * (ch @ y)
            elif e.left:
                du[dst, 0] += a * (c @ x)
                du[dst, 1] += a * (c @ y + cw @ x)
            else:
                du[dst, 0] += a * (x @ ch)
                du[dst, 1] += a * (y @ ch)
        return du

    u0 = np.zeros((spec.channels, 2, s.D, s.D), np.complex128)
    if adj:
        u0[-1, 1] = np.eye(s.D)
    else:
        u0[0, 0] = s.rho
        u0[0, 1] = rw
    return _integrate(rhs, u0, s, t_eval=list(points))


def excitation_observable_action(ctx, w, v, rw, spec):
    """src/excitation_action.jl:103-126."""
    s = ctx.state
    points, weights = _quadrature(s.n_quad, s.max_point)
    fwd, fend = action_trajectory(ctx, w, v, rw, spec, points, False)
    bwd, bend = action_trajectory(ctx, w, v, rw, spec, points, True)
    npts = len(points)
    if len(fwd) != npts or len(bwd) != npts:
        raise ExcitationFailure("missing trajectory samples")
    out = np.zeros_like(w)
    wrho = w @ s.rho
    for i in range(npts):
        x, a = fwd[i], bwd[npts - 1 - i]
        dx, j = weights[i] * 2 * abs(points[i]), weights[i] * transformed_source(points[i])
        for k in range(spec.channels):
            x0, x1, a0, a1 = x[k, 0], x[k, 1], a[k, 0], a[k, 1]
            out += dx * (bra_transfer_pull(s, a0, x0) + bra_transfer_pull(s, a1, x1) + a1 @ w @ x0 -
                         np.trace(a1 @ x0) * wrho)
        for e in spec.edges:
            if e.left:
                continue
            b = a[e.dst, 0] @ x[e.src, 0] + a[e.dst, 1] @ x[e.src, 1]
            out += (j * e.coefficient) * bra_source_pull(ctx, b, spec.commutator)
    out += action_boundaries(ctx, w, v, rw, bend[0, 0], bend[0, 1], fend[-1, 0])
    return out


def _action_resolve(ctx, b, z, shifted=False, adj=False):
    return _solve_block(ctx, b, _resolvent_factor(ctx, z, 1 if shifted else 0), adj)


def excitation_kinetic_action_sum(ctx, w, v, rw, terms, n):
    """src/excitation_action.jl:135-179."""
    s = ctx.state
    out = np.zeros_like(w)
    wrho = w @ s.rho
    ta0, ta1, tx0 = (np.zeros_like(w) for _ in range(3))
    prepared = [(coef, spec, *action_operators(ctx, w, v, spec.commutator)) for coef, spec in terms]
    for theta, nw in zip(*kinetic_grid(n)):
        z = math.cos(theta)
        for coef, spec, c, cw in prepared:
            orders = ((False, True), (True, False)) if spec.kinetic else \
                ((False, False), (False, True), (True, False), (True, True))
            weight = nw * coef / 2
            ch = c.conj().T
            for left, secondleft in orders:
                b0 = c @ s.rho if left else s.rho @ ch
                b1 = c @ rw + cw @ s.rho if left else rw @ ch
                x0 = _action_resolve(ctx, b0, z)
                x1 = _action_resolve(ctx, b1 + z * ket_insertion(s, v, w, x0), z, shifted=True)
                xout = c @ x0 if secondleft else x0 @ ch
                if not secondleft:
                    a0, a1 = np.zeros_like(w), ch.copy()
                    out += weight * bra_source_pull(ctx, x1, spec.commutator)
                else:
                    a0, a1 = cw.copy(), c.copy()
                a1 = _action_resolve(ctx, a1, z, adj=True)
                a0 = _action_resolve(ctx, a0 + z * ket_trace_dual(s, v, w, a1), z, shifted=True, adj=True)
                out += (weight * z) * (bra_transfer_pull(s, a0, x0) + bra_transfer_pull(s, a1, x1) + a1 @ w @ x0 -
                                       np.trace(a1 @ x0) * wrho)
                if left:
                    a0, a1 = a0 @ c + a1 @ cw, a1 @ c
                else:
                    out += weight * bra_source_pull(ctx, a0 @ s.rho + a1 @ rw, spec.commutator)
                    a0, a1 = ch @ a0, ch @ a1
                ta0 += weight * a0
                ta1 += weight * a1
                tx0 += weight * xout
    out += action_boundaries(ctx, w, v, rw, ta0, ta1, tx0)
    return out


def excitation_kinetic_action(ctx, w, v, rw, terms, o=KineticOptions()):
    n = o.minnodes
    previous = excitation_kinetic_action_sum(ctx, w, v, rw, terms, n)
    while 2 * n <= o.maxnodes:
        n *= 2
        current = excitation_kinetic_action_sum(ctx, w, v, rw, terms, n)
        if np.linalg.norm(current - previous) <= o.atol + o.rtol * np.linalg.norm(current):
            return current
        previous = current
    raise ExcitationFailure(f"excitation kinetic action quadrature did not converge by {n} nodes")


def _is_quadratic(spec):
    return spec.channels == 3 and not spec.kinetic and spec.edges == monomial_spec(2).edges


def hamiltonian_action(ctx, w, terms, kinetic_backend="resolvent", kinetic_nodes=None,
                       kinetic_options=KineticOptions(), parts=False):
    """excitation_hamiltonian_action (src/excitation_action.jl:194-228).  With parts=True returns
    (kinetic part, potential part)."""
    s = ctx.state
    w = np.asarray(w, np.complex128)
    if w.shape != s.R.shape:
        raise ValueError("tangent dimensions do not match the state")
    v = -s.R.conj().T @ w
    rw = excitation_poisson(ctx, ket_insertion(s, v, w, s.rho), 1)
    resolved = lambda spec: spec.kinetic or _is_quadratic(spec)
    kinetics = [t for t in terms if resolved(t[1])]
    potentials = [t for t in terms if kinetic_backend == "ode" or not resolved(t[1])]
    kin = np.zeros_like(w)
    if kinetic_backend == "resolvent" and kinetics:
        kin = (excitation_kinetic_action(ctx, w, v, rw, kinetics, kinetic_options) if kinetic_nodes is None
               else excitation_kinetic_action_sum(ctx, w, v, rw, kinetics, kinetic_nodes))
    pot = np.zeros_like(w)
    for coef, spec in potentials:
        pot = pot + coef * excitation_observable_action(ctx, w, v, rw, spec)
    return (kin, pot) if parts else kin + pot


# --------------------------------------------------------------------------------------------
# white coordinates of one parity sector (src/excitation_eigsolve.jl:5-130)
# --------------------------------------------------------------------------------------------
@dataclass
class Coordinates:
    state: XState
    rotation: np.ndarray
    whitening: np.ndarray
    indices: list
    parity: object
    sector: str

    def __len__(self):
        return len(self.indices)

    def tangent_matrix(self, x):
        d = self.state.D
        m = np.zeros((d, d), np.complex128)
        for (i, j), xi in zip(self.indices, x):
            m[i, j] = xi
        return m @ self.whitening.conj().T

    def tangent_coordinates(self, w):
        m = w @ self.state.rho @ self.whitening
        return np.array([m[i, j] for i, j in self.indices])

    def covector_coordinates(self, f):
        m = f @ self.whitening
        return np.array([m[i, j] for i, j in self.indices])


def balanced_parity(s: XState, tol=1e-12):
    d = s.D
    if d % 2:
        return None
    z = np.concatenate([np.ones(d // 2), -np.ones(d // 2)])
    zz = np.outer(z, z)
    K = 1j * (s.Q + 0.5 * s.R.conj().T @ s.R)        # Q = -iK - R^+R/2
    err = max(np.linalg.norm(zz * s.R + s.R) / max(1, np.linalg.norm(s.R)),
              np.linalg.norm(zz * K - K) / max(1, np.linalg.norm(K)), np.linalg.norm(zz * s.rho - s.rho))
    return z if err <= tol else None


def coordinates(s: XState, sector="auto", gauge="density", metric_tol=0.0):
    """TangentCoordinates (src/excitation_eigsolve.jl:60-112) with the balanced parity of a Z2 state."""
    z = balanced_parity(s)
    if sector == "auto":
        sector = "all" if z is None else "odd"
    d = s.D
    blocks = [list(range(d))] if z is None else [list(np.where(z == 1)[0]), list(np.where(z == -1)[0])]
    u = np.zeros((d, d), np.complex128)
    lam = np.zeros(d)
    for b in blocks:
        vals, vecs = np.linalg.eigh(s.rho[np.ix_(b, b)])
        u[np.ix_(b, b)] = vecs
        lam[b] = vals
    cutoff = max(metric_tol, 10 * d * np.finfo(float).eps) * lam.max()
    keep = lam > cutoff
    if gauge == "density":
        rho = u.conj().T @ s.rho @ u
        t = XState(u.conj().T @ s.R @ u, u.conj().T @ s.Q @ u, (rho + rho.conj().T) / 2, s.n_quad, s.max_point,
                   s.atol, s.rtol)
        rotation = u
    else:
        t, rotation = s, np.eye(d, dtype=np.complex128)
    B = np.zeros((d, d), np.complex128)
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
    allowed = lambda i, j: sector == "all" or (sector == "odd") == (z[i] == z[j])
    indices = [(i, j) for j in columns for i in range(d) if allowed(i, j)]
    return Coordinates(t, rotation, B, indices, z, sector)


def sector_matrix(c: Coordinates, apply):
    """Dense sector operator: column j is covector_coordinates(F(W(e_j))) (the assembly=:action
    matrix of src/excitations.jl:160-172)."""
    n = len(c)
    h = np.zeros((n, n), np.complex128)
    for j in range(n):
        e = np.zeros(n, np.complex128)
        e[j] = 1
        h[:, j] = c.covector_coordinates(apply(c.tangent_matrix(e)))
    return h
