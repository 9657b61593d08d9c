"""The infinite-line kinetic backend on the GPU (matrix-free Krylov, kinetic_backend="krylov")
against the oracle's dense-Schur resolvent, through the C ABI.

Bars as in the reference's own tests (test/kinetic_krylov.jl, test/runtests.jl:261-290): the
coherent closed forms; Krylov against dense Schur at the same node count, energy 1e-11 and
finished gradient 1e-8; fg_phi4 with the Krylov kinetic terms against the oracle's resolvent
fg_phi4 (energy 1e-11, gradient 1e-8) and against the wide-window ODE path (1e-10, 2e-8); every
shifted solve accepted on its true relative residual; failures raise.
"""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rg():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import rcmps_gpu
    return rcmps_gpu


def test_coherent_closed_forms(rg):
    note("case: kinetic krylov coherent D=1")
    s = rg.State(np.array([[0.3 + 0.1j]]), np.zeros((1, 1)))
    assert rg.kinetic1(s, gradient=False) == pytest.approx(0.1, abs=1e-14)
    v2, g2 = rg.kinetic2(s)
    assert v2 == 0.0
    assert np.abs(g2).max() == 0.0


@pytest.mark.parametrize("d,seed,scaling", [(3, 90, 0.2), (4, 7, 0.3), (8, 314, 0.3)])
def test_krylov_matches_dense_resolvent(rg, d, seed, scaling):
    note(f"case: kinetic krylov vs dense D={d} seed={seed}")
    r, k = rg.random_state_arrays(d, scaling=scaling, seed=seed)
    s = rg.State(r, k)
    (v, G), st = rg.kinetic1(s, return_stats=True)
    v2, G2 = rg.kinetic2(s)
    ref_v, ref_c, ref_e, used = oracle.kinetic_resolvent(s.R, s.Q, s.rho, metric_reg=s.manifold_param.metric_reg)
    note(f"  nodes={st['kinetic_nodes']} (oracle {used}) basis={st['kinetic_basis']} "
         f"actions={st['kinetic_actions']} residual={st['kinetic_residual']:.2e}")
    assert v + v2 == pytest.approx(ref_v, abs=1e-11, rel=1e-11)
    assert st["kinetic_residual"] <= 1e-12
    # the combined gradient against the oracle's finished gradient of the same covector/environment
    Gref = oracle.metric_solve_full(s.R, s.Q, s.rho, ref_c, ref_e, s.manifold_param.metric_reg)
    assert np.abs(G + G2 - Gref).max() <= 1e-8 * max(1.0, np.abs(Gref).max())
    s.close()


def test_fg_phi4_krylov_against_oracle_and_wide_window(rg):
    note("case: fg_phi4 kinetic krylov D=3 seed 90")
    r, k = rg.random_state_arrays(3, scaling=0.2, seed=90)
    s = rg.State(r, k)
    ip = s.integrator_param
    ip.n_quad, ip.max_point, ip.atol, ip.rtol = 256, 30.0, 1e-13, 1e-12
    e, G = rg.fg_phi4(s, 2.0, kinetic_backend="krylov")
    kw = dict(n_quad=256, max_point=30.0, atol=1e-13, rtol=1e-12)
    ref = oracle.fg_phi4(s.R, s.Q, s.rho, 2.0, kinetic_backend="resolvent", **kw)
    assert e == pytest.approx(ref["energy"], abs=1e-11)
    np.testing.assert_allclose(G, ref["gradient"], atol=1e-8, rtol=1e-8)
    e_ode, G_ode = rg.fg_phi4(s, 2.0)
    assert e == pytest.approx(e_ode, abs=1e-10)
    np.testing.assert_allclose(G, G_ode, atol=2e-8, rtol=2e-8)
    s.close()


def test_failures(rg):
    note("case: kinetic krylov failure modes")
    r, k = rg.random_state_arrays(8, scaling=0.3, seed=314)
    s = rg.State(r, k)
    with pytest.raises(rg.SolverFailure):
        rg.kinetic1(s, krylov_options=rg.KineticKrylovOptions(initialdim=1, maxdim=1, tolerance=1e-13))
    with pytest.raises(rg.SolverFailure):
        rg.kinetic2(s, kinetic_options=rg.KineticOptions(minnodes=2, maxnodes=4, atol=1e-15, rtol=0.0))
    with pytest.raises(ValueError):
        rg.kinetic1(s, krylov_options=rg.KineticKrylovOptions(tolerance=float("inf")))
    with pytest.raises(ValueError):
        rg.kinetic1(s, kinetic_options=rg.KineticOptions(minnodes=1))
    # the context's controls are restored to the defaults by the next call
    v = rg.kinetic1(s, gradient=False)
    assert np.isfinite(v)
    s.close()


def test_bases_follow_the_state_size(rg):
    """Regression: the context's cached Arnoldi bases are re-sized when a state of another padded
    size follows (they were once reused at the old vector length)."""
    note("case: kinetic krylov D=8 then D=40 then D=8 on one context")
    vals = []
    for d in (8, 40, 8):
        r, k = rg.random_state_arrays(d, scaling=0.3, seed=314)
        s = rg.State(r, k)
        vals.append(rg.kinetic1(s, gradient=False))
        s.close()
    assert vals[0] == vals[2] and np.isfinite(vals[1])


def test_d64_krylov_kinetic(rg):
    import time
    note("case: fg_phi4 kinetic krylov D=64 g=2 seed 42")
    r, k = rg.random_state_arrays(64, scaling=0.2, seed=42)
    s = rg.State(r, k)
    t0 = time.perf_counter()
    e, G, st = rg.fg_phi4(s, 2.0, kinetic_backend="krylov", return_stats=True)
    dt = time.perf_counter() - t0
    e_ode, G_ode = rg.fg_phi4(s, 2.0)
    note(f"  E krylov={e!r} E ode={e_ode!r} diff={abs(e - e_ode):.2e} |dG|={np.abs(G - G_ode).max():.2e} "
         f"time={dt:.3f}s nodes={st['kinetic_nodes']} basis={st['kinetic_basis']} residual={st['kinetic_residual']:.1e}")
    assert np.isfinite(e) and abs(e - e_ode) < 1e-6
    s.close()
