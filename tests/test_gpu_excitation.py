"""The tangent-space excitation action and eigensolver on the GPU.

Parity with the oracle (oracle/excitation.py, pinned by closed forms and the reference's golden
levels): on seeded D = 2 and D = 4 states at p = 0 and 0.7 the kinetic part (fixed 32-node theta
grid) and the :phi^4: part of F(W) agree with the oracle's dense-Schur and SciPy-DOP853 values within
1e-9 relative to |F| (both integrate to 1e-12); the Poisson solves agree within 1e-11.  Closed forms:
vacuum sqrt(1 + p^2) (1e-11), coherent D = 1 at g = 2 omega + 24 0.3^2 / omega (1e-8).  Batching:
three tangents in one call equal three single calls to 1e-12.  Connected element: the device pair
route agrees with the oracle's kinetic sum (1e-10) and :phi^4: ODE (1e-9), with <u, F(w)> (2e-8), is
Hermitian between (u, w) and (w, u) (1e-9), and its adaptive grid agrees with the oracle's 512-node
value (1e-9).  Golden levels: excitation_eigsolve on the reference's D = 4 Z2 state (tests/golden)
reproduces the recorded Davidson Ritz energies of both sectors at p = 0 and 0.4, all four validation
energies (standard and refined fresh actions, standard and refined bilinear elements), the validated
energy and the acceptance, within their tolerance 1e-11.
"""
import json
import math
import time
import os
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle
from oracle import excitation as X

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).parent / "golden"


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rg():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import rcmps_gpu
    return rcmps_gpu


def random_arrays(d, seed, scaling):
    rng = np.random.default_rng(seed)
    x = scaling * (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d)))
    K = (x + x.conj().T) / 2
    R = scaling * (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d)))
    return R, K


def oracle_state(s, ip):
    return X.XState(s.R, s.Q, s.rho, n_quad=ip.n_quad, max_point=ip.max_point, atol=ip.atol, rtol=ip.rtol)


@pytest.mark.parametrize("dense", [True, False], ids=["dense_transfer", "krylov"])
@pytest.mark.parametrize("d,seed,p", [(2, 42, 0.0), (2, 42, 0.7), (4, 7, 0.7)])
def test_action_parts_match_oracle(rg, d, seed, p, dense):
    note(f"case: excitation action D={d} seed={seed} p={p} ({'dense transfer' if dense else 'Krylov'} resolvents)")
    from rcmps_gpu import IntegratorParameters
    from rcmps_gpu.excitation import ExcitationContext, ExcitationModel, action_batch
    R, K = random_arrays(d, seed, 0.5)
    s = rg.State(R, K, steady_tol=1e-13)
    s.integrator_param = IntegratorParameters(n_quad=64, max_point=18.0, atol=1e-12, rtol=1e-12)
    ctx = ExcitationContext(s, p, dense_transfer=dense)
    assert (getattr(s, "_transfer_cond", None) is not None) == dense
    rng = np.random.default_rng(seed + 1)
    w = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    model = ExcitationModel("phi4", 1.0)
    (kin,), st = action_batch(ctx, [w], model, 32, 1, return_stats=True)
    (pot,), st2 = action_batch(ctx, [w], model, 32, 2, return_stats=True)
    note(f"  kinetic bases {st['kinetic_basis'][:2]} actions {st['kinetic_actions']} residual {st['kinetic_residual']:.2e};"
         f" ODE attempts {st2['attempts'][4]} accepted {st2['accepted'][4]}")
    xs = oracle_state(s, s.integrator_param)
    octx = X.context(xs, p)
    okin, opot = X.hamiltonian_action(octx, w, X.phi4_terms(1.0), kinetic_nodes=32, parts=True)
    ek = np.abs(kin - okin).max() / np.abs(okin).max()
    ep = np.abs(pot - opot).max() / np.abs(opot).max()
    note(f"  kinetic rel {ek:.2e}  potential rel {ep:.2e}")
    assert ek < 1e-9
    assert ep < 1e-9
    full = rg.excitation_hamiltonian_action(ctx, w, model=model, kinetic_nodes=32)
    assert np.abs(full - (kin + pot)).max() <= 1e-12 * np.abs(full).max()


@pytest.mark.parametrize("name,D", [("phi4_g2_D16_z2_state", 16), ("phi4_g2.7_D32_z2_state", 32)])
def test_dense_transfer_resolvents_match_krylov(rg, name, D):
    """The kinetic part of F(W) for a batch of three tangents at two momenta on the reference's Z2
    checkpoints: resolvents from the transfer's eigendecomposition against the Krylov bases (solved
    to 1e-11), within 1e-8 of |F|; and the batch (its node sums made for all tangents together)
    against each tangent alone with the dense resolvents, within 1e-12 of |F| (the pinned Poisson
    products of the set-up take all right-hand sides at once, so the bits may differ)."""
    note(f"case: dense transfer vs Krylov resolvents, {name}")
    from rcmps_gpu.excitation import ExcitationContext, ExcitationModel, action_batch
    from rcmps_gpu.jld2 import read_jld2
    rec = read_jld2(GOLDEN / f"{name}.jld2")
    s = rg.State(rec.R, rec.K, steady_tol=1e-12)
    model = ExcitationModel("phi4", rec.coupling)
    rng = np.random.default_rng(5)
    ws = [rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D)) for _ in range(3)]
    tight = rg.KineticKrylovOptions(tolerance=1e-11, fallback_tolerance=1e-11)
    out = {}
    for dense in (True, False):
        ctxs = [ExcitationContext(s, p, dense_transfer=dense) for p in (0.1, 0.1, 0.6)]
        t0 = time.perf_counter()
        fs = action_batch(ctxs, ws, model, 64, parts=1, krylov_options=None if dense else tight)
        out[dense] = (fs, time.perf_counter() - t0)
        if dense:
            for c, w, f in zip(ctxs, ws, fs):
                (alone,) = action_batch(c, [w], model, 64, parts=1)
                assert np.abs(alone - f).max() <= 1e-12 * np.abs(f).max()
        for c in ctxs:
            c.close()
    note(f"  kinetic part of 3 tangents: dense {out[True][1]:.3f} s, Krylov {out[False][1]:.3f} s, "
         f"condition {getattr(s, '_transfer_cond', None)}")
    for a, b in zip(out[True][0], out[False][0]):
        assert np.abs(a - b).max() <= 1e-8 * np.abs(b).max()
    s.close()


@pytest.mark.parametrize("nodes", [112, 130])
def test_krylov_kinetic_batch_equals_single_tangents(rg, nodes):
    """The kinetic part of F(W) on the Krylov route (D = 40, padded to 48) for two tangents at two
    momenta in one call against each tangent alone, within 1e-8 of |F| (the lockstep bases of a batch
    may take more steps than one tangent's; observed: the same bits).  Each tangent's node sums are one segment of a segmented
    reduction in chunks of ceil(nodes / 56) nodes at this size: 112 nodes split evenly, 130 do not and
    take one reduction per tangent."""
    note(f"case: Krylov-route kinetic batch vs single tangents, D=40, {nodes} nodes")
    from rcmps_gpu.excitation import ExcitationContext, ExcitationModel, action_batch
    r, k = rg.random_state_arrays(40, scaling=0.2, seed=7)
    s = rg.State(r, k, steady_tol=1e-12)
    model = ExcitationModel("phi4", 2.0)
    rng = np.random.default_rng(nodes)
    ws = [rng.standard_normal((40, 40)) + 1j * rng.standard_normal((40, 40)) for _ in range(2)]
    tight = rg.KineticKrylovOptions(tolerance=1e-11, fallback_tolerance=1e-11)
    ctxs = [ExcitationContext(s, p) for p in (0.1, 0.6)]
    t0 = time.perf_counter()
    fs = action_batch(ctxs, ws, model, nodes, parts=1, krylov_options=tight)
    note(f"  batch of 2: {time.perf_counter() - t0:.3f} s")
    for c, w, f in zip(ctxs, ws, fs):
        assert np.isfinite(f).all()
        (alone,) = action_batch(c, [w], model, nodes, parts=1, krylov_options=tight)
        note(f"  |batch - alone| / |F| = {np.abs(alone - f).max() / np.abs(f).max():.2e}")
        assert np.abs(alone - f).max() <= 1e-8 * np.abs(f).max()
    for c in ctxs:
        c.close()
    s.close()


@pytest.mark.parametrize("d,seed,p", [(2, 42, 0.7), (4, 7, 0.0)])
def test_connected_element_matches_oracle(rg, d, seed, p):
    """<Phi_p(u)|H - E0|Phi_p(w)> by the device pair environments, four-block ODE and kinetic Krylov
    sum against the oracle's dense route (both integrate to 1e-12), and against <u, F(w)>."""
    note(f"case: excitation element D={d} seed={seed} p={p}")
    from rcmps_gpu import IntegratorParameters
    from rcmps_gpu.excitation import ExcitationContext, ExcitationModel, action_batch, excitation_element
    R, K = random_arrays(d, seed, 0.5)
    s = rg.State(R, K, steady_tol=1e-13)
    s.integrator_param = IntegratorParameters(n_quad=64, max_point=18.0, atol=1e-12, rtol=1e-12)
    ctx = ExcitationContext(s, p)
    rng = np.random.default_rng(seed + 2)
    u = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    w = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    model = ExcitationModel("phi4", 1.0)
    (kin,) = excitation_element(ctx, [u], [w], model=model, kinetic_nodes=32, parts=1)
    (pot,) = excitation_element(ctx, [u], [w], model=model, kinetic_nodes=32, parts=2)
    xs = oracle_state(s, s.integrator_param)
    pr = X.pair(X.context(xs, p), u, w)
    terms = X.phi4_terms(1.0)
    okin = X.excitation_kinetic_sum(pr, [t for t in terms if t[1].kinetic], 32)
    opot = 1.0 * X.excitation_observable(pr, X.monomial_spec(4))
    note(f"  kinetic {kin} vs {okin}; potential {pot} vs {opot}")
    assert abs(kin - okin) <= 1e-10 * max(1.0, abs(okin))
    assert abs(pot - opot) <= 1e-9 * max(1.0, abs(opot))
    (f,) = action_batch(ctx, [w], model, 32)
    assert abs(np.vdot(u, f) - (kin + pot)) <= 2e-8 * max(1.0, abs(kin + pot))


def test_matrix_element_hermitian_and_adaptive(rg):
    """excitation_matrix_element(s, p, u, w) with the adaptive theta grid: the sesquilinear form is
    Hermitian (<u|H|w> = conj <w|H|u> from two independent pair integrations), its diagonal is real,
    and it agrees with the oracle's element on a fixed fine grid to the quadrature tolerance."""
    note("case: excitation_matrix_element D=4 seed=5 p=0.3 adaptive")
    from rcmps_gpu import IntegratorParameters
    from rcmps_gpu.excitation import ExcitationModel, excitation_matrix_element
    R, K = random_arrays(4, 5, 0.5)
    s = rg.State(R, K, steady_tol=1e-13)
    s.integrator_param = IntegratorParameters(n_quad=64, max_point=18.0, atol=1e-12, rtol=1e-12)
    rng = np.random.default_rng(11)
    u = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    w = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    model = ExcitationModel("phi4", 1.0)
    uw, wu, uu = excitation_matrix_element(s, 0.3, [u, w, u], [w, u, u], model=model)
    single = excitation_matrix_element(s, 0.3, u, w, model=model)
    note(f"  <u|H|w> {uw}, conj <w|H|u> {np.conj(wu)}, <u|H|u> {uu}")
    scale = max(1.0, abs(uw))
    assert isinstance(single, complex) and abs(single - uw) <= 1e-12 * scale
    assert abs(uw - np.conj(wu)) <= 1e-9 * scale
    assert abs(uu.imag) <= 1e-9 * max(1.0, abs(uu))
    xs = oracle_state(s, s.integrator_param)
    pr = X.pair(X.context(xs, 0.3), u, w)
    terms = X.phi4_terms(1.0)
    ref = X.excitation_kinetic_sum(pr, [t for t in terms if t[1].kinetic], 512) \
        + X.excitation_observable(pr, X.monomial_spec(4))
    assert abs(uw - ref) <= 1e-9 * scale


@pytest.mark.parametrize("dense", [True, False], ids=["dense_transfer", "gmres"])
def test_poisson_solves_match_oracle(rg, dense):
    """The pinned Poisson solves, by the transfer's eigenbasis (dense_transfer) or by GMRES."""
    note(f"case: excitation Poisson solves D=4 p=0.7 ({'dense transfer' if dense else 'GMRES'})")
    from rcmps_gpu.excitation import ExcitationContext, action_dual_poisson, excitation_poisson
    R, K = random_arrays(4, 3, 0.5)
    s = rg.State(R, K, steady_tol=1e-13)
    ctx = ExcitationContext(s, 0.7, dense_transfer=dense)
    octx = X.context(oracle_state(s, s.integrator_param), 0.7)
    rng = np.random.default_rng(9)
    b = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    for shift in (-1, 0, 1):
        x = excitation_poisson(ctx, b, shift)
        assert np.abs(x - X.excitation_poisson(octx, b, shift)).max() < 1e-11
    for shifted in (False, True):
        x = action_dual_poisson(ctx, b, shifted)
        assert np.abs(x - X.action_dual_poisson(octx, b, shifted)).max() < 1e-11


@pytest.mark.parametrize("p", [0.0, 0.7])
def test_closed_forms(rg, p):
    note(f"case: excitation closed forms p={p}")
    from rcmps_gpu import IntegratorParameters
    from rcmps_gpu.excitation import ExcitationContext, ExcitationModel
    vac = rg.State(np.zeros((1, 1)), np.zeros((1, 1)))
    a = rg.excitation_hamiltonian_action(ExcitationContext(vac, p), np.ones((1, 1)), model=ExcitationModel("phi4", 0.0))
    assert abs(a[0, 0] - math.sqrt(1 + p * p)) < 1e-11
    r = 0.3 + 0.1j
    coh = rg.State(np.array([[r]]), np.zeros((1, 1)))
    coh.integrator_param = IntegratorParameters(n_quad=128, max_point=30.0)
    a = rg.excitation_hamiltonian_action(ExcitationContext(coh, p), np.ones((1, 1)), model=ExcitationModel("phi4", 2.0))
    omega = math.sqrt(1 + p * p)
    assert abs(a[0, 0] - (omega + 24 * 0.3 ** 2 / omega)) < 1e-8


def test_batched_tangents_equal_single_calls(rg):
    note("case: excitation action batch of 3 D=4")
    from rcmps_gpu import IntegratorParameters
    from rcmps_gpu.excitation import ExcitationContext, ExcitationModel, action_batch
    R, K = random_arrays(4, 11, 0.5)
    s = rg.State(R, K)
    s.integrator_param = IntegratorParameters(n_quad=32, max_point=14.0, atol=1e-12, rtol=1e-12)
    ctx = ExcitationContext(s, 0.4)
    rng = np.random.default_rng(5)
    ws = [rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)) for _ in range(3)]
    model = ExcitationModel("phi4", 1.0)
    together = action_batch(ctx, ws, model, 32)
    for w, f in zip(ws, together):
        alone = action_batch(ctx, [w], model, 32)[0]
        assert np.abs(alone - f).max() <= 1e-10 * np.abs(f).max()


def d4_state(rg):
    from rcmps_gpu import IntegratorParameters
    from rcmps_gpu.jld2 import read_jld2
    rec = read_jld2(GOLDEN / "phi4_g1_D4_z2_state.jld2")
    s = rg.State(rec.R, rec.K, steady_tol=1e-13)
    s.integrator_param = IntegratorParameters(n_quad=32, max_point=14.0, atol=1e-13, rtol=1e-12)
    s.model_param.name = "phi4"
    s.model_param.coupling = rec.coupling
    return s


@pytest.mark.parametrize("sector,p", [("odd", 0.0), ("odd", 0.4), ("even", 0.0), ("even", 0.4)])
def test_d4_eigsolve_matches_golden_levels(rg, sector, p):
    note(f"case: excitation_eigsolve D=4 {sector} p={p}")
    from rcmps_gpu import DavidsonOptions
    levels = json.loads((GOLDEN / "excitation_references.json").read_text())["levels"]
    gold = sorted((e["level"], e) for e in levels
                  if e["state"] == "phi4_g1_D4" and e["sector"] == sector and e["momentum"] == p)
    s = d4_state(rg)
    opts = DavidsonOptions(tol=1e-8, maxops=60, maxspace=8, retain=2, refresh=True)
    r = rg.excitation_eigsolve(s, p, nev=2, sector=sector, kinetic_nodes=32, davidson_options=opts)
    note(f"  energies {r['energies']} numops {r['numops']} start {r['start']} stop {r['stop_reason']}")
    assert r["converged"] and r["start"] == gold[0][1]["start"]
    for (level, e), energy, v in zip(gold, r["energies"], r["validation"]):
        assert abs(energy - e["ritz_energy"]) <= e["ritz_tolerance"], (level, energy, e["ritz_energy"])
        four = (v["standard_energy"], v["refined_energy"], v["bilinear_energy"], v["refined_bilinear_energy"])
        note(f"  level {level}: four energies {four} (golden {e['four_energies']})")
        for mine, theirs in zip(four, e["four_energies"]):
            assert abs(mine - theirs) <= e["tolerance"]
        assert abs(r["validated_energies"][level - 1] - e["validated_energy"]) <= e["tolerance"]
        assert v["accepted"] == e["accepted"]


def test_d16_eigsolve_matches_golden_level(rg):
    """The optimized D = 16 Z2 ground state (g = 2) at the settings of the reference's GPU comparison
    runs: window 30, n_quad 512, tolerances 1e-13 / 1e-12, 256 kinetic nodes, thick restarts."""
    note("case: excitation_eigsolve D=16 odd p=0 (gpu_comparison settings)")
    from rcmps_gpu import DavidsonOptions, IntegratorParameters
    from rcmps_gpu.jld2 import read_jld2
    levels = json.loads((GOLDEN / "excitation_references.json").read_text())["levels"]
    e = [x for x in levels if x["name"] == "D16_odd_p0"][0]
    rec = read_jld2(GOLDEN / "phi4_g2_D16_z2_state.jld2")
    s = rg.State(rec.R, rec.K, steady_tol=1e-12)
    s.integrator_param = IntegratorParameters(n_quad=512, max_point=30.0, atol=1e-13, rtol=1e-12)
    s.model_param.name = "phi4"
    s.model_param.coupling = rec.coupling
    opts = DavidsonOptions(tol=1e-7, maxops=200, maxspace=128, retain=12, refresh=False)
    r = rg.excitation_eigsolve(s, 0.0, nev=1, kinetic_nodes=256, davidson_options=opts)
    v = r["validation"][0]
    note(f"  energy {r['energies'][0]!r} (golden {e['ritz_energy']!r}) numops {r['numops']} (golden {e['numops']}) "
         f"refined {v['refined_energy']!r} (golden {e['four_energies'][1]!r}) timings {r['timings']}")
    assert r["converged"] and r["sector"] == "odd" and r["dimension"] == e["dimension"]
    assert abs(r["energies"][0] - e["ritz_energy"]) <= e["ritz_tolerance"]
    assert abs(v["refined_energy"] - e["four_energies"][1]) <= e["ritz_tolerance"]


@pytest.mark.parametrize("name", ["D32_odd_p0.05", "D32_even_p0"])
def test_d32_eigsolve_matches_golden_level(rg, name):
    """The optimized D = 32 Z2 ground state (g = 2.7) at the GPU-comparison settings."""
    note(f"case: excitation_eigsolve {name} (gpu_comparison settings)")
    from rcmps_gpu import DavidsonOptions, IntegratorParameters
    from rcmps_gpu.jld2 import read_jld2
    levels = json.loads((GOLDEN / "excitation_references.json").read_text())["levels"]
    e = [x for x in levels if x["name"] == name][0]
    rec = read_jld2(GOLDEN / "phi4_g2.7_D32_z2_state.jld2")
    s = rg.State(rec.R, rec.K, steady_tol=1e-12)
    s.integrator_param = IntegratorParameters(n_quad=512, max_point=30.0, atol=1e-13, rtol=1e-12)
    s.model_param.name = "phi4"
    s.model_param.coupling = rec.coupling
    opts = DavidsonOptions(tol=1e-7, maxops=200, maxspace=160, retain=12, refresh=False)
    r = rg.excitation_eigsolve(s, e["momentum"], nev=1, sector=e["sector"], kinetic_nodes=256, davidson_options=opts)
    v = r["validation"][0]
    note(f"  energy {r['energies'][0]!r} (golden {e['ritz_energy']!r}) numops {r['numops']} (golden {e['numops']}) "
         f"refined {v['refined_energy']!r} (golden {e['four_energies'][1]!r}) timings {r['timings']}")
    assert r["converged"] and r["dimension"] == e["dimension"]
    assert abs(r["energies"][0] - e["ritz_energy"]) <= e["ritz_tolerance"]
    assert abs(v["refined_energy"] - e["four_energies"][1]) <= e["ritz_tolerance"]


def test_d4_lockstep_momenta_match_golden_levels(rg):
    """The Davidson searches of p = 0 and 0.4 in lockstep (one device call per round applies one
    action per momentum) give the golden levels of both momenta."""
    note("case: iterative_dispersion lockstep D=4 odd p in (0, 0.4)")
    from rcmps_gpu import DavidsonOptions, iterative_dispersion
    levels = json.loads((GOLDEN / "excitation_references.json").read_text())["levels"]
    s = d4_state(rg)
    opts = DavidsonOptions(tol=1e-8, maxops=60, maxspace=8, retain=2, refresh=True)
    rs = iterative_dispersion(s, [0.0, 0.4], nev=2, sector="odd", kinetic_nodes=32, davidson_options=opts,
                              lockstep=True)
    for r, p in zip(rs, (0.0, 0.4)):
        gold = sorted((e["level"], e) for e in levels
                      if e["state"] == "phi4_g1_D4" and e["sector"] == "odd" and e["momentum"] == p)
        note(f"  p={p} energies {r['energies']} numops {r['numops']} rounds {r['device_rounds']}")
        assert r["numops"] == gold[0][1]["numops"]
        for (level, e), energy in zip(gold, r["energies"]):
            assert abs(energy - e["ritz_energy"]) <= e["ritz_tolerance"]
    assert rs[0]["device_rounds"] <= max(r["numops"] for r in rs)


def test_repeated_actions_release_their_workspaces(rg):
    """Every per-call device buffer of the action is released (a leak once ran a long dispersion
    solve out of memory)."""
    note("case: excitation action memory after repeated calls D=16")
    import torch
    from rcmps_gpu import IntegratorParameters
    from rcmps_gpu.excitation import ExcitationContext, ExcitationModel, action_batch
    R, K = random_arrays(16, 3, 0.3)
    s = rg.State(R, K)
    s.integrator_param = IntegratorParameters(n_quad=32, max_point=14.0)
    ctx = ExcitationContext(s, 0.3)
    rng = np.random.default_rng(4)
    ws = [rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16)) for _ in range(4)]
    model = ExcitationModel("phi4", 1.0)
    action_batch(ctx, ws, model, 32)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(8):
        action_batch(ctx, ws, model, 32)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    note(f"  free memory change {(free1 - free0) / 2**20:.1f} MiB")
    assert free0 - free1 < 64 * 2**20


def test_nine_momentum_dispersion_lockstep(rg):
    """The Z2 dispersion workload: RandomStateZ2(32) (scaling 0.2, seed 42), g = 2, odd sector, nev = 2
    at p = 0, 0.25, ..., 2 with the nine Davidson searches in lockstep (one device call applies one
    action per momentum) at the reference's default solver settings.  Properties: the sector has
    dimension 512; at every momentum the two Ritz levels are ordered; every accepted level has its four
    validation energies within the acceptance spread; most momenta accept both levels; the lowest level
    rises from p = 0 to p = 2; and the p = 0 search equals a standalone excitation_eigsolve at p = 0
    (same start, same operator count, energies to 1e-10)."""
    note("case: dispersion RandomStateZ2(32) g=2 odd nev=2 p=0:0.25:2 lockstep")
    from rcmps_gpu import IntegratorParameters, iterative_dispersion
    r, k = rg.random_state_z2_arrays(32, scaling=0.2, seed=42)
    s = rg.State(r, k)
    s.integrator_param = IntegratorParameters(n_quad=128, max_point=20.0, atol=1e-11, rtol=1e-10)
    s.model_param.name = "phi4"
    s.model_param.coupling = 2.0
    momenta = [0.25 * i for i in range(9)]
    rs = iterative_dispersion(s, momenta, nev=2, sector="odd", lockstep=True, strict=False)
    lowest = []
    accepted_both = 0
    for p, res in zip(momenta, rs):
        note(f"  p={p}: energies {res['energies']} numops {res['numops']} accepted {res['accepted']} "
             f"stop {res['stop_reason']}")
        assert res["dimension"] == 512 and len(res["energies"]) == 2
        assert res["energies"][0] <= res["energies"][1] + 1e-12
        for v in res["validation"]:
            if v["accepted"]:
                assert v["energy_spread"] <= res["validation_tol"]
        accepted_both += bool(res["accepted"])
        lowest.append(res["energies"][0])
    assert accepted_both >= 7
    assert lowest[-1] > lowest[0]
    alone = rg.excitation_eigsolve(s, 0.0, nev=2, sector="odd")
    note(f"  standalone p=0: energies {alone['energies']} numops {alone['numops']}")
    assert alone["numops"] == rs[0]["numops"]
    np.testing.assert_allclose(alone["energies"], rs[0]["energies"], rtol=0, atol=1e-10)
    s.close()
