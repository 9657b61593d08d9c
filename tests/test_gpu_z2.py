"""Z2-symmetric states in the trajectory kernel and the action ODE, through the C ABI.

For a state with R odd and K, rho0 even under the split of the basis into halves, every channel of
the transfer ODEs has a definite parity, so half of each channel's parity blocks stay exactly zero
and a tile product touches one parity block of k-tiles or none.  When the halves are whole 16 x 16
tiles (D = 32, 64) the kernel skips the others (the default); RCMPS_FLAG_DENSE runs every k-tile.
The skipped k-steps multiply exact zeros, so both give the same bits: energies, gradients and
step counts are compared exactly, in both precisions, on the reference's D = 16 (halves inside one
tile: nothing to skip) and D = 32 checkpoints and on a seeded RandomStateZ2(64).  The same holds for
the excitation action's :phi^4: ODE (RCMPS_EXC_DENSE) when every tangent of a call has one parity
(odd, even; a mixed call runs every k-tile).  The times of both are printed.
"""
import sys

import numpy as np
import pytest

from conftest import FIXTURES

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


def fixture_state(rg, name, n_quad):
    from rcmps_gpu.jld2 import read_jld2
    rec = read_jld2(FIXTURES / f"{name}.jld2")
    s = rg.State(rec.R, rec.K, steady_tol=1e-12)
    s.integrator_param.n_quad = n_quad
    s.integrator_param.atol, s.integrator_param.rtol = 1e-13, 1e-12
    return s, rec.coupling


@pytest.mark.parametrize("case", ["D16_fixture", "D32_fixture", "D64_random"])
def test_z2_blocks_give_the_same_bits(rg, case):
    note(f"case: Z2 blocks vs dense, {case}")
    if case == "D16_fixture":
        s, g = fixture_state(rg, "phi4_g2_D16_z2_state", 192)
    elif case == "D32_fixture":
        s, g = fixture_state(rg, "phi4_g2.7_D32_z2_state", 192)
    else:
        r, k = rg.random_state_z2_arrays(64, scaling=0.2, seed=42)
        s, g = rg.State(r, k), 2.0
        s.integrator_param.n_quad = 256
    for precision in ("fp64", "split"):
        e1, g1, st1 = rg.fg_phi4(s, g, return_stats=True, precision=precision)
        e0, g0, st0 = rg.fg_phi4(s, g, return_stats=True, precision=precision, z2_blocks=False)
        note(f"  {precision}: E = {e1:.15g}, trajectory {st1['trajectory_ms']:.2f} ms with the parity blocks, "
             f"{st0['trajectory_ms']:.2f} ms dense, attempts {list(st1['attempts'])}")
        assert e1 == e0
        assert np.array_equal(g1, g0)
        assert list(st1["attempts"]) == list(st0["attempts"])
        assert list(st1["rejected"]) == list(st0["rejected"])
    s.close()


@pytest.mark.parametrize("name,D", [("phi4_g2.7_D32_z2_state", 32), ("phi4_g2_D16_z2_state", 16)])
def test_action_ode_z2_blocks_give_the_same_bits(rg, name, D):
    """The :phi^4: action ODE of odd and even tangents (and a mixed pair, which runs every k-tile) on
    the reference's Z2 checkpoints: the same bits with and without the parity blocks."""
    note(f"case: action ODE Z2 blocks vs dense, {name}")
    from rcmps_gpu import IntegratorParameters
    from rcmps_gpu.excitation import ExcitationContext, ExcitationModel, action_batch
    from rcmps_gpu.jld2 import read_jld2
    rec = read_jld2(FIXTURES / f"{name}.jld2")
    s = rg.State(rec.R, rec.K, steady_tol=1e-12)
    s.integrator_param = IntegratorParameters(n_quad=64, max_point=20.0, atol=1e-11, rtol=1e-10)
    ctx = ExcitationContext(s, 0.3)
    model = ExcitationModel("phi4", rec.coupling)
    rng = np.random.default_rng(11)
    h = D // 2

    def tangent(parity):
        w = rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D))
        if parity == 1:
            w[:h, :h] = 0
            w[h:, h:] = 0
        elif parity == 0:
            w[:h, h:] = 0
            w[h:, :h] = 0
        return w

    for label, ws in (("odd", [tangent(1), tangent(1)]), ("even", [tangent(0)]), ("mixed", [tangent(1), tangent(0)])):
        f1, st1 = action_batch(ctx, ws, model, 32, parts=2, return_stats=True)
        f0, st0 = action_batch(ctx, ws, model, 32, parts=2, return_stats=True, z2_blocks=False)
        note(f"  {label}: ODE {st1['trajectory_ms']:.1f} ms with the parity blocks, {st0['trajectory_ms']:.1f} ms dense, "
             f"attempts {st1['attempts'][4]}")
        for a, b in zip(f1, f0):
            assert np.array_equal(a, b)
        assert st1["attempts"][4] == st0["attempts"][4]
    ctx.close()
    s.close()
