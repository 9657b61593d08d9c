"""Batched fg_phi4: many states of uneven D and coupling in persistent launches (one per padded
size), through the C ABI.

Each system is integrated by the same task code whatever else shares the launch, so the batched
energies and gradients equal the single-state calls bit for bit; the small states are also
checked against the oracle (energy 1e-12, gradient 1e-10 relative).  The mixed batch follows the
distribution of the 256-state workload (D in {8, 16, 24, 32, 48, 64}, g uniform in [0.5, 4],
complex normal entries scaled 0.2, seeded), at 24 states.
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


def mixed_states(rg, n, seed0, n_quad=64):
    rng = np.random.default_rng(2024)
    dims = rng.choice([8, 16, 24, 32, 48, 64], size=n)
    gs = rng.uniform(0.5, 4.0, size=n)
    states = []
    for i, d in enumerate(dims):
        r, k = rg.random_state_arrays(int(d), scaling=0.2, seed=seed0 + i)
        s = rg.State(r, k)
        s.integrator_param.n_quad = n_quad
        states.append(s)
    return states, [float(g) for g in gs]


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_batch_equals_single_calls(rg, precision):
    note(f"case: batch of 9 small/uneven states, {precision}")
    dims = [3, 8, 16, 5, 20, 32, 12, 40, 16]
    states = []
    for i, d in enumerate(dims):
        r, k = rg.random_state_arrays(d, scaling=0.3, seed=100 + i)
        s = rg.State(r, k)
        s.integrator_param.n_quad = 32
        states.append(s)
    gs = [0.5 + 0.4 * i for i in range(len(dims))]
    E, G, stats = rg.fg_phi4_batch(states, gs, precision=precision, return_stats=True)
    for i, s in enumerate(states):
        e1, g1 = rg.fg_phi4(s, gs[i], precision=precision)
        assert E[i] == e1, (i, E[i], e1)
        assert np.array_equal(G[i], g1), i
        assert stats[i]["status"] == 0
    if precision == "fp64":
        for i in (0, 1, 3):
            s = states[i]
            ref = oracle.fg_phi4(s.R, s.Q, s.rho, gs[i], n_quad=32)
            assert E[i] == pytest.approx(ref["energy"], rel=1e-12)
            assert np.abs(G[i] - ref["gradient"]).max() <= 1e-10 * np.abs(ref["gradient"]).max()
    for s in states:
        s.close()


def test_mixed_batch_workload_distribution(rg):
    note("case: mixed batch of 24 states, D in {8..64}, g in [0.5, 4]")
    states, gs = mixed_states(rg, 24, 0)
    E, G, stats = rg.fg_phi4_batch(states, gs, return_stats=True)
    Es, _ = rg.fg_phi4_batch(states, gs, gradient=False)
    for i, s in enumerate(states):
        assert np.isfinite(E[i]) and G[i].shape == (s.D, s.D)
        assert Es[i] == E[i]          # the energy-only launch integrates the same forward systems
    # one single-state call per distinct D as a spot check
    seen = set()
    for i, s in enumerate(states):
        if s.D in seen:
            continue
        seen.add(s.D)
        e1, g1 = rg.fg_phi4(s, gs[i])
        assert E[i] == e1 and np.array_equal(G[i], g1)
    note(f"  trajectory ms per launch class: {sorted({round(x['trajectory_ms'], 2) for x in stats})}")
    for s in states:
        s.close()
