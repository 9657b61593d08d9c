"""The :phi^4: action ODE in one persistent launch against the replayed launch graphs, through the C ABI.

The action ODE (excitation_observable_action, src/excitation_action.jl:56-126) runs either as one
cooperative launch of rc_quartic_ode (RCMPS_EXC_PERSISTENT), every tile's workgroup resident and the
DOP853 controller run redundantly in each, or as a captured graph of fourteen launches per attempt
replayed by the host (RCMPS_EXC_GRAPH).  Both run the same arithmetic in the same order and take the
same step decisions, so the actions and the step counts must be the same bits: checked on the
reference's Z2 checkpoints
(D = 16, 32: one and two tangents, with and without the parity blocks) and on seeded random states
(D = 8, 24: one and three tangents, the latter with one tile of padding).  The times of both are
printed.
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


def state(rg, case):
    from rcmps_gpu import IntegratorParameters
    from rcmps_gpu.jld2 import read_jld2
    if case.startswith("phi4"):
        rec = read_jld2(FIXTURES / f"{case}.jld2")
        s, g = rg.State(rec.R, rec.K, steady_tol=1e-12), rec.coupling
    else:
        d = int(case[len("random_D"):])
        r, k = rg.random_state_arrays(d, scaling=0.3, seed=d)
        s, g = rg.State(r, k, steady_tol=1e-12), 2.0
    s.integrator_param = IntegratorParameters(n_quad=64, max_point=20.0, atol=1e-11, rtol=1e-10)
    return s, g


@pytest.mark.parametrize("case,ntan", [("phi4_g2_D16_z2_state", 1), ("phi4_g2.7_D32_z2_state", 2),
                                       ("random_D8", 3), ("random_D24", 1), ("random_D24", 3)])
def test_persistent_launch_gives_the_graph_bits(rg, case, ntan):
    note(f"case: persistent action ODE vs graphs, {case}, {ntan} tangent(s)")
    from rcmps_gpu.excitation import ExcitationContext, ExcitationModel, action_batch
    s, g = state(rg, case)
    ctx = ExcitationContext(s, 0.4)
    model = ExcitationModel("phi4", g)
    rng = np.random.default_rng(5)
    D, h = s.D, s.D // 2
    ws = []
    for _ in range(ntan):
        w = rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D))
        if case.startswith("phi4"):   # odd tangents: the parity blocks apply
            w[:h, :h] = 0
            w[h:, h:] = 0
        ws.append(w)
    for z2 in ((True, False) if case.startswith("phi4") else (True,)):
        f1, st1 = action_batch(ctx, ws, model, 32, parts=2, return_stats=True, z2_blocks=z2, persistent=True)
        f0, st0 = action_batch(ctx, ws, model, 32, parts=2, return_stats=True, z2_blocks=z2, persistent=False)
        note(f"  z2 blocks {z2}: ODE {st1['trajectory_ms']:.1f} ms persistent, {st0['trajectory_ms']:.1f} ms graphs, "
             f"attempts {st1['attempts'][4]}, rejected {st1['rejected'][4]}")
        ph = st1["phase_us"][4]
        note(f"  persistent workgroup 0: products {ph[0] / 1e3:.1f} ms, exchanges {ph[1] / 1e3:.1f} ms, "
             f"controller {ph[2] / 1e3:.1f} ms, launch {ph[3] / 1e3:.1f} ms")
        assert st1["attempts"][4] == st0["attempts"][4] > 0
        assert st1["rejected"][4] == st0["rejected"][4]
        for a, b in zip(f1, f0):
            assert np.isfinite(a).all()
            assert np.array_equal(a, b)
    ctx.close()
    s.close()


def test_step_limit_stops_both_routes_alike(rg):
    """A step budget far below what the integration needs ends both routes with MAX_STEPS (the
    controller's own stop, taken by every workgroup of the persistent launch at the same attempt),
    and the context gives the same bits afterwards."""
    note("case: action ODE step limit, D=16 checkpoint, both routes")
    from rcmps_gpu import IntegratorParameters
    from rcmps_gpu._lib import DeviceControlFailure
    from rcmps_gpu.excitation import ExcitationContext, ExcitationModel, action_batch
    s, g = state(rg, "phi4_g2_D16_z2_state")
    ctx = ExcitationContext(s, 0.4)
    model = ExcitationModel("phi4", g)
    w = np.random.default_rng(9).standard_normal((16, 16)) + 0j
    (before,) = action_batch(ctx, [w], model, 32, parts=2, persistent=True)
    ip = ctx.integrator
    ctx.integrator = IntegratorParameters(n_quad=ip.n_quad, max_point=ip.max_point, atol=ip.atol, rtol=ip.rtol,
                                          maxiters=20)
    for persistent in (True, False):
        with pytest.raises(DeviceControlFailure) as exc:
            action_batch(ctx, [w], model, 32, parts=2, persistent=persistent)
        note(f"  persistent={persistent}: {exc.value}")
        assert exc.value.status == 1
    ctx.integrator = ip
    (after,) = action_batch(ctx, [w], model, 32, parts=2, persistent=True)
    assert np.array_equal(before, after)
    ctx.close()
    s.close()
