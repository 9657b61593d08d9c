"""Host logic of the excitation solver (rcmps_gpu/excitation.py), on the CPU.

- davidson (src/davidson.jl:106-206) on seeded Hermitian matrices: the lowest nev eigenvalues of
  numpy's eigh within the residual target, with fresh restarts and with thick restarts, the
  residuals recomputed from the stored actions, a preconditioned search needing fewer operator
  calls than an unpreconditioned one, :breakdown in a spanned space, the Hermiticity guard, and the
  argument checks of check_davidson (src/davidson.jl:31-52).
- relativistic_diagnostics (src/excitations.jl:250-277) on the exact dispersion sqrt(1 + p^2):
  c^2 = 1, intercept 1 (test/excitations.jl:57-63).
"""
import math

import numpy as np
import pytest

from rcmps_gpu.excitation import DavidsonOptions, check_davidson, davidson, relativistic_diagnostics
from rcmps_gpu import SolverFailure


def hermitian(n, seed, spread=10.0):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    vals = np.sort(rng.uniform(0, spread, n))
    return (q * vals) @ q.conj().T, vals


@pytest.mark.parametrize("refresh", [True, False])
def test_davidson_finds_the_lowest_levels(refresh):
    A, vals = hermitian(60, 3)
    rng = np.random.default_rng(1)
    x0 = rng.standard_normal(60) + 1j * rng.standard_normal(60)
    opts = DavidsonOptions(tol=1e-9, maxops=400, maxspace=12, retain=3, refresh=refresh)
    r = davidson(lambda x: A @ x, x0, nev=2, options=opts)
    assert r["converged"] and r["stop_reason"] == "converged"
    assert np.allclose(r["energies"], vals[:2], atol=1e-12)
    for v, a, res in zip(r["vectors"], r["actions"], r["residuals"]):
        assert abs(np.linalg.norm(v) - 1) < 1e-12
        assert np.linalg.norm(A @ v - a) < 1e-10
        assert res <= 1e-9
    assert r["numops"] == len(r["history"])


def test_preconditioning_reduces_operator_calls():
    rng = np.random.default_rng(5)
    P = 0.5 * (rng.standard_normal((80, 80)) + 1j * rng.standard_normal((80, 80)))
    A = np.diag(np.linspace(1.0, 100.0, 80)) + (P + P.conj().T) / 2    # diagonally dominant
    vals = np.linalg.eigvalsh(A)
    d = np.real(np.diag(A))
    x0 = np.zeros(80, complex)
    x0[np.argmin(d)] = 1
    opts = DavidsonOptions(tol=1e-8, maxops=400, maxspace=20, retain=2, refresh=False)
    plain = davidson(lambda x: A @ x, x0, nev=1, options=opts)
    pre = davidson(lambda x: A @ x, x0, nev=1, options=opts,
                   precondition=lambda r: r / np.maximum(d - plain["energies"][0], 1.0))
    assert plain["converged"] and pre["converged"]
    assert abs(pre["energies"][0] - vals[0]) < 1e-10
    assert pre["numops"] < plain["numops"]


def test_batched_start_block_equals_sequential():
    A, _ = hermitian(30, 7)
    rng = np.random.default_rng(2)
    starts = [rng.standard_normal(30) + 1j * rng.standard_normal(30) for _ in range(3)]
    opts = DavidsonOptions(tol=1e-10, maxops=100, maxspace=10, retain=3)
    a = davidson(lambda x: A @ x, starts, nev=3, options=opts)
    b = davidson(lambda x: A @ x, starts, nev=3, options=opts, batch_op=lambda xs: [A @ x for x in xs])
    assert a["numops"] == b["numops"]
    assert np.allclose(a["energies"], b["energies"], atol=1e-12)


def test_breakdown_and_guards():
    A = np.diag([1.0, 2.0, 3.0]).astype(complex)
    # the start vectors span the whole space: no new direction, residual exactly representable
    r = davidson(lambda x: A @ x, [np.eye(3)[0] + 0.3 * np.eye(3)[1], np.eye(3)[2]], nev=1,
                 options=DavidsonOptions(tol=1e-30, maxops=10, maxspace=4, retain=1, stall_window=0))
    assert r["stop_reason"] in ("breakdown", "converged")
    skew = A + 1e-3j * np.array([[0, 1, 0], [1, 0, 0], [0, 0, 0]])
    with pytest.raises(SolverFailure):
        davidson(lambda x: skew @ x, [np.ones(3), np.array([1.0, -1, 0])], nev=1,
                 options=DavidsonOptions(maxspace=4, retain=1))
    with pytest.raises(ValueError):
        check_davidson(DavidsonOptions(maxspace=2), 2)
    with pytest.raises(ValueError):
        check_davidson(DavidsonOptions(tol=0.0), 1)
    with pytest.raises(ValueError):
        davidson(lambda x: A @ x, np.zeros(3), nev=1)


def test_relativistic_diagnostics_of_the_free_dispersion():
    ps = np.array([-2.0, -0.5, 0.0, 0.5, 2.0])
    e = np.sqrt(1 + ps ** 2)[:, None]
    d = relativistic_diagnostics(dict(momenta=ps, energies=e))
    assert abs(d["c_squared"] - 1) < 1e-11 and abs(d["intercept"] - 1) < 1e-11
    assert d["invariant_spread"] < 1e-10 and d["fit_residual"] < 1e-12
    scaled = relativistic_diagnostics(dict(momenta=ps, energies=2 * e), target_speed=2)
    assert np.allclose(scaled["invariant_mass_squared"], 4.0, atol=1e-10)
    with pytest.raises(ValueError):
        relativistic_diagnostics(dict(momenta=np.array([-1.0, 1.0]), energies=np.ones((2, 1))))
    with pytest.raises(ValueError):
        relativistic_diagnostics(dict(momenta=np.array([0.0, 1.0]), energies=-np.ones((2, 1))))


def test_kinetic_quadrature_controls_are_checked():
    """The element's adaptive theta grid rejects controls with maxnodes < 2 minnodes, minnodes < 2 or
    a non-positive atol (src/excitations.jl:89-91), and fixed grids below two nodes."""
    from rcmps_gpu import KineticOptions
    from rcmps_gpu.excitation import _check_kinetic_options, check_kinetic_nodes
    _check_kinetic_options(KineticOptions())
    for bad in (KineticOptions(minnodes=1, maxnodes=8), KineticOptions(minnodes=32, maxnodes=48),
                KineticOptions(atol=0.0), KineticOptions(rtol=-1e-3), KineticOptions(atol=math.inf)):
        with pytest.raises(ValueError):
            _check_kinetic_options(bad)
    assert check_kinetic_nodes(2) == 2
    with pytest.raises(ValueError):
        check_kinetic_nodes(1)
