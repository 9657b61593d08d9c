"""ctypes wrapper of the CPU oracle (oracle/oracle.c).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
import this package, and only as the checker.  Matrices are complex128, row-major (C order).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

STATUS = {0: "success", 1: "max_steps", 2: "step_underflow", 3: "nonfinite_error",
          5: "missing_samples", 10: "solver_failure", 11: "bad_argument"}


class Integrator(ctypes.Structure):
    _fields_ = [("n_quad", ctypes.c_int), ("max_point", ctypes.c_double), ("atol", ctypes.c_double),
                ("rtol", ctypes.c_double), ("initial_step", ctypes.c_double),
                ("max_steps", ctypes.c_long)]


class OracleFailure(RuntimeError):
    pass


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        p = ctypes.c_void_p
        _lib.or_source.restype = ctypes.c_double
        _lib.or_source.argtypes = [ctypes.c_double]
        _lib.or_quadrature.argtypes = [ctypes.c_int, ctypes.c_double, p, p]
        _lib.or_fg_phi4.argtypes = [ctypes.c_int, p, p, p, ctypes.c_double, ctypes.POINTER(Integrator),
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                    p, p, p, p, p, p]
        _lib.or_fg_phi4_rk.argtypes = [ctypes.c_int, p, p, p, ctypes.c_double, ctypes.POINTER(Integrator),
                                       ctypes.c_double, ctypes.c_int, ctypes.c_int, p, p, p]
        _lib.or_trajectory.argtypes = [ctypes.c_int, p, p, p, ctypes.c_int, ctypes.POINTER(Integrator), p, p]
        _lib.or_kinetic_resolvent.argtypes = [ctypes.c_int, p, p, p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                              p, p, p, p]
        _lib.or_system_len.restype = ctypes.c_size_t
        _lib.or_system_len.argtypes = [ctypes.c_int, ctypes.c_int]
        _lib.or_steady_state.argtypes = [ctypes.c_int, p, p, ctypes.c_double, ctypes.c_int, p, p]
        _lib.or_boundary.argtypes = [ctypes.c_int, p, p, p, p, ctypes.c_double, ctypes.c_int, p]
        _lib.or_nojump.argtypes = [ctypes.c_int, p, p, p, ctypes.c_int, p]
        _lib.or_metric_solve.argtypes = [ctypes.c_int, p, ctypes.c_double, p, p]
        _lib.or_lindblad.argtypes = [ctypes.c_int, p, p, p, p, ctypes.c_int]
        _lib.or_rhs_full.argtypes = [ctypes.c_int, p, p, ctypes.c_int, ctypes.c_int, ctypes.c_double, p, p]
        _lib.or_schur.argtypes = [ctypes.c_int, p, p]
    return _lib


def _c(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a, a.ctypes.data_as(ctypes.c_void_p)


def _check(status, what):
    if status != 0:
        raise OracleFailure(f"{what} failed: {STATUS.get(status, status)}")


def source(t: float) -> float:
    return lib().or_source(float(t))


def quadrature(n_quad: int, max_point: float):
    pts = np.empty(2 * n_quad)
    wts = np.empty(2 * n_quad)
    lib().or_quadrature(int(n_quad), float(max_point), pts.ctypes.data_as(ctypes.c_void_p),
                        wts.ctypes.data_as(ctypes.c_void_p))
    return pts, wts


def integrator(n_quad=128, max_point=20.0, atol=1e-11, rtol=1e-10, initial_step=0.01,
               max_steps=1_000_000):
    return Integrator(int(n_quad), float(max_point), float(atol), float(rtol), float(initial_step),
                      int(max_steps))


def canonical_q(R, K):
    R = np.asarray(R, dtype=np.complex128)
    K = np.asarray(K, dtype=np.complex128)
    K = (K + K.conj().T) / 2
    return -1j * K - R.conj().T @ R / 2


def fg_phi4(R, Q, rho, g, *, n_quad=128, max_point=20.0, atol=1e-11, rtol=1e-10, initial_step=0.01,
            max_steps=1_000_000, metric_reg=1e-3, boundary_tol=1e-13, boundary=True,
            boundary_method="iterative", nthreads=6, kinetic_backend="ode"):
    """E, G and diagnostics of the oracle's fg_phi4 (kinetic_backend "ode" or "resolvent")."""
    d = len(R)
    R, pr = _c(R)
    Q, pq = _c(Q)
    rho, prho = _c(rho)
    ip = integrator(n_quad, max_point, atol, rtol, initial_step, max_steps)
    e = ctypes.c_double()
    G = np.zeros((d, d), np.complex128)
    cov = np.zeros((d, d), np.complex128)
    env = np.zeros((d, d), np.complex128)
    terms = np.zeros(3)
    stats = np.zeros(30, dtype=np.int64)
    flags = (0 if boundary else 1) | (2 if boundary_method == "dense" else 0) | (4 if kinetic_backend == "resolvent" else 0)
    st = lib().or_fg_phi4(d, pr, pq, prho, float(g), ctypes.byref(ip), float(metric_reg),
                          float(boundary_tol), flags, int(nthreads), ctypes.addressof(e),
                          G.ctypes.data, cov.ctypes.data, env.ctypes.data, terms.ctypes.data,
                          stats.ctypes.data)
    _check(st, "oracle fg_phi4")
    return dict(energy=e.value, gradient=G, covector=cov, environment=env, term_values=terms,
                stats=stats.reshape(6, 5))


def fg_phi4_rk(R, K, g, *, rho=None, n_quad=128, max_point=20.0, atol=1e-11, rtol=1e-10, initial_step=0.01,
               max_steps=1_000_000, boundary_tol=1e-13, boundary=True, boundary_method="iterative",
               steady_tol=1e-12, nthreads=6):
    """(energy, grad_R, grad_K) of the state (R, K): dE = Re sum(conj(grad_R) dR + conj(grad_K) dK)
    for Hermitian dK.  rho0 is solved by the oracle unless given."""
    d = len(R)
    Q = canonical_q(R, K)
    if rho is None:
        rho = steady_state(Q, R, tol=steady_tol)
    R, pr = _c(R)
    Q, pq = _c(Q)
    rho, prho = _c(rho)
    ip = integrator(n_quad, max_point, atol, rtol, initial_step, max_steps)
    e = ctypes.c_double()
    GR = np.zeros((d, d), np.complex128)
    GK = np.zeros((d, d), np.complex128)
    flags = (0 if boundary else 1) | (2 if boundary_method == "dense" else 0)
    st = lib().or_fg_phi4_rk(d, pr, pq, prho, float(g), ctypes.byref(ip), float(boundary_tol), flags,
                             int(nthreads), ctypes.addressof(e), GR.ctypes.data, GK.ctypes.data)
    _check(st, "oracle fg_phi4_rk")
    return e.value, GR, GK


def kinetic_resolvent(R, Q, rho, *, minnodes=32, maxnodes=512, atol=1e-11, rtol=1e-10, metric_reg=1e-3,
                      boundary_tol=1e-13, gradient=True):
    """kinetic1 + kinetic2 on the infinite line by the dense-Schur angular resolvent with node
    doubling: (value, covector, environment, nodes used)."""
    d = len(R)
    R, pr = _c(R)
    Q, pq = _c(Q)
    rho, prho = _c(rho)
    v = ctypes.c_double()
    used = ctypes.c_int()
    cov = np.zeros((d, d), np.complex128)
    env = np.zeros((d, d), np.complex128)
    st = lib().or_kinetic_resolvent(d, pr, pq, prho, int(minnodes), int(maxnodes), float(atol), float(rtol),
                                    float(metric_reg), float(boundary_tol), 1 if gradient else 0,
                                    ctypes.addressof(v), cov.ctypes.data, env.ctypes.data, ctypes.addressof(used))
    _check(st, "oracle kinetic_resolvent")
    return v.value, cov, env, used.value


def metric_solve_full(R, Q, rho, covector, environment, metric_reg=1e-3, boundary_tol=1e-13):
    """finish_gradient (src/observables.jl:201-212) of a raw covector/environment pair."""
    R = np.asarray(R, np.complex128)
    h = boundary_environment(Q, R, rho, environment, tol=boundary_tol, method="dense")
    c = np.asarray(covector, np.complex128) + rho @ R.conj().T @ h - rho @ h @ R.conj().T
    return _metric_solve(rho, metric_reg, c)


def _metric_solve(rho, reg, cov):
    d = len(rho)
    rho, prho = _c(rho)
    cov, pc = _c(cov)
    G = np.zeros((d, d), np.complex128)
    _check(lib().or_metric_solve(d, prho, float(reg), pc, G.ctypes.data), "oracle metric_solve")
    return G


def system_len(d: int, system: int) -> int:
    return int(lib().or_system_len(int(d), int(system)))


def trajectory(R, Q, rho, system, *, n_quad=128, max_point=20.0, atol=1e-11, rtol=1e-10,
               initial_step=0.01, max_steps=1_000_000):
    """Reduced samples (2 n_quad + 1, len) of one system (2*term + adjoint) and its stats."""
    d = len(R)
    R, pr = _c(R)
    Q, pq = _c(Q)
    rho, prho = _c(rho)
    ip = integrator(n_quad, max_point, atol, rtol, initial_step, max_steps)
    length = system_len(d, system)
    samples = np.zeros((2 * n_quad + 1, length), np.complex128)
    stats = np.zeros(5, dtype=np.int64)
    st = lib().or_trajectory(d, pq, pr, prho, int(system), ctypes.byref(ip), samples.ctypes.data,
                             stats.ctypes.data)
    _check(st, "oracle trajectory")
    return samples, stats


def steady_state(Q, R, *, tol=1e-11, method="cptp"):
    d = len(Q)
    Q, pq = _c(Q)
    R, pr = _c(R)
    rho = np.zeros((d, d), np.complex128)
    it = ctypes.c_int()
    st = lib().or_steady_state(d, pq, pr, float(tol), 1 if method == "dense" else 0, rho.ctypes.data,
                               ctypes.addressof(it))
    _check(st, "oracle steady_state")
    return rho


def boundary_environment(Q, R, rho, a, *, tol=1e-11, method="iterative"):
    d = len(Q)
    Q, pq = _c(Q)
    R, pr = _c(R)
    rho, prho = _c(rho)
    a, pa = _c(a)
    h = np.zeros((d, d), np.complex128)
    st = lib().or_boundary(d, pq, pr, prho, pa, float(tol), 1 if method == "dense" else 0, h.ctypes.data)
    _check(st, "oracle boundary")
    return h


def nojump(Q, R, Y, *, adjoint=False):
    d = len(Q)
    Q, pq = _c(Q)
    R, pr = _c(R)
    Y, py = _c(Y)
    out = np.zeros((d, d), np.complex128)
    _check(lib().or_nojump(d, pq, pr, py, 1 if adjoint else 0, out.ctypes.data), "oracle nojump")
    return out


def schur(A):
    d = len(A)
    T = np.ascontiguousarray(A, dtype=np.complex128).copy()
    Z = np.zeros((d, d), np.complex128)
    _check(lib().or_schur(d, T.ctypes.data, Z.ctypes.data), "oracle schur")
    return T, Z


def lindblad(Q, R, X, *, adjoint=False):
    d = len(Q)
    Q, pq = _c(Q)
    R, pr = _c(R)
    X, px = _c(X)
    out = np.zeros((d, d), np.complex128)
    lib().or_lindblad(d, pq, pr, px, out.ctypes.data, 1 if adjoint else 0)
    return out


def rhs_full(Q, R, term, adj, t, U):
    d = len(Q)
    Q, pq = _c(Q)
    R, pr = _c(R)
    U, pu = _c(U)
    dU = np.zeros_like(U)
    lib().or_rhs_full(d, pq, pr, int(term), int(adj), float(t), pu, dU.ctypes.data)
    return dU


def metric(rho, x, y, metric_reg=1e-3):
    """src/optimization.jl:1: 2 Re tr(x^dagger y (rho + reg I))."""
    d = len(rho)
    return float(2 * np.vdot(x, y @ (rho + metric_reg * np.eye(d))).real)


def retraction(R, K, W, alpha):
    """src/optimization.jl:7-13: (R + alpha W, K + i alpha/2 (W^dagger R - R^dagger W))."""
    R, K, W = (np.asarray(x, dtype=np.complex128) for x in (R, K, W))
    k = K + 1j * alpha / 2 * (W.conj().T @ R - R.conj().T @ W)
    return R + alpha * W, (k + k.conj().T) / 2
