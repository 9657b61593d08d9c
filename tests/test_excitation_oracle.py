"""The excitation oracle (oracle/excitation.py) pinned before it checks the GPU action.

- Closed forms: the vacuum dispersion sqrt(1 + p^2) and the coherent D = 1 state at g = 2,
  omega + 24 0.3^2 / omega, for both the reverse-contracted action and the bilinear element
  (the reference's bars, test/excitations.jl:57-69 and test/excitation_action.jl:24: 1e-11, 1e-8).
- Consistency of two independent routes: <u, F(w)> equals the connected matrix element
  <Phi_p(u)|H - E0|Phi_p(w)> of the four-block ODE and resolvent (test/excitation_action.jl:9-11:
  2e-8 absolute), and the pair environments solve the connected transfer equation
  (test/excitations.jl:81-88).
- The reference's golden levels (tests/golden/excitation_references.json from
  test/fixtures/excitation_references.toml): the lowest two eigenvalues of the dense odd and
  even sector operators of the D = 4 Z2 state at p = 0 and 0.4 equal the recorded Davidson Ritz
  energies within their stated tolerance 1e-11 (observed within 1.3e-12).
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from oracle import excitation as X

GOLDEN = Path(__file__).parent / "golden"


def random_state(d, seed, scaling):
    rng = np.random.default_rng(seed)
    x = scaling * (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d)))
    K = (x + x.conj().T) / 2
    R = scaling * (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d)))
    Q = -1j * K - 0.5 * R.conj().T @ R
    return R, Q, oracle.steady_state(Q, R, tol=1e-14)


@pytest.mark.parametrize("p", [0.0, 0.7])
def test_vacuum_dispersion(p):
    s = X.XState(np.zeros((1, 1), complex), np.zeros((1, 1), complex), np.ones((1, 1), complex))
    a = X.hamiltonian_action(X.context(s, p), np.ones((1, 1), complex), X.phi4_terms(0.0))
    assert abs(a[0, 0] - math.sqrt(1 + p * p)) < 1e-11


@pytest.mark.parametrize("p", [0.0, 0.7, 2.0])
def test_coherent_state_closed_form(p):
    r = 0.3 + 0.1j
    s = X.XState(np.array([[r]]), np.array([[-abs(r) ** 2 / 2]], complex), np.ones((1, 1), complex),
                 n_quad=128, max_point=30.0)
    ctx = X.context(s, p)
    omega = math.sqrt(1 + p * p)
    exact = omega + 24 * 0.3 ** 2 / omega
    w = np.ones((1, 1), complex)
    assert abs(X.hamiltonian_action(ctx, w, X.phi4_terms(2.0))[0, 0] - exact) < 1e-8
    assert abs(X.excitation_element(X.pair(ctx, w, w), X.phi4_terms(2.0)) - exact) < 1e-8


@pytest.mark.parametrize("p", [0.0, 0.7])
def test_action_contracts_to_the_matrix_element(p):
    R, Q, rho = random_state(2, 42, 0.5)
    s = X.XState(R, Q, rho, n_quad=96, max_point=24.0, atol=1e-12, rtol=1e-12)
    rng = np.random.default_rng(1)
    u = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    w = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    ctx = X.context(s, p)
    terms = X.phi4_terms(1.0)
    a = X.hamiltonian_action(ctx, w, terms)
    exact = X.excitation_element(X.pair(ctx, u, w), terms)
    assert abs(np.vdot(u, a) - exact) < 2e-8 + 2e-9 * abs(exact)
    # linearity of the action in w (complex scalar)
    assert np.abs(X.hamiltonian_action(ctx, (1 + 1j) * w, terms) - (1 + 1j) * a).max() < 2e-8


def test_pair_environments_are_stationary():
    R, Q, rho = random_state(2, 82, 0.7)
    s = X.XState(R, Q, rho)
    rng = np.random.default_rng(84)
    u = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    w = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    for p in (0.0, 0.6):
        pr = X.pair(X.context(s, p), u, w)
        assert np.linalg.norm(X.excitation_transfer(pr.right, pr)) < 1e-10
        assert abs(X.excitation_contract(pr, pr.right)) < 1e-11
        x = rng.standard_normal((4, 2, 2)) + 1j * rng.standard_normal((4, 2, 2))
        assert abs(X.excitation_contract(pr, X.excitation_transfer(x, pr))) < 1e-10
        assert abs(pr.overlap - np.trace(w @ rho @ u.conj().T)) < 1e-14
        assert abs(np.trace(pr.right[1])) + abs(np.trace(pr.right[2])) + abs(np.trace(pr.right[3])) < 1e-10


def d4_state():
    from rcmps_gpu.jld2 import read_jld2
    rec = read_jld2(GOLDEN / "phi4_g1_D4_z2_state.jld2")
    Q = -1j * rec.K - 0.5 * rec.R.conj().T @ rec.R
    rho = oracle.steady_state(Q, rec.R, tol=1e-14)
    return X.XState(rec.R, Q, rho, n_quad=32, max_point=14.0, atol=1e-13, rtol=1e-12), rec.coupling


@pytest.mark.parametrize("sector,p", [("odd", 0.0), ("even", 0.4)])
def test_d4_sector_spectrum_matches_golden_ritz_energies(sector, p):
    levels = json.loads((GOLDEN / "excitation_references.json").read_text())["levels"]
    gold = sorted((e["level"], e["ritz_energy"], e["ritz_tolerance"]) for e in levels
                  if e["state"] == "phi4_g1_D4" and e["sector"] == sector and e["momentum"] == p)
    assert len(gold) == 2
    s, g = d4_state()
    c = X.coordinates(s, sector=sector)
    assert len(c) == 8
    ctx = X.context(c.state, p)
    h = X.sector_matrix(c, lambda w: X.hamiltonian_action(ctx, w, X.phi4_terms(g), kinetic_nodes=32))
    assert np.linalg.norm(h - h.conj().T) < 1e-9
    ev = np.linalg.eigvalsh((h + h.conj().T) / 2)
    for (level, energy, tol), e in zip(gold, ev):
        assert abs(e - energy) <= tol, (level, e, energy)
