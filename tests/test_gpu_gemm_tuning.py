"""The batched-product autotuner: every tile variant gives the same bits, and the tuner's table
holds a valid choice with its measured times.

Each variant is forced in turn on the same computations that run batched and single products on
the device (fg_phi4 pullback and boundary solve at D = 16 and 40; the excitation action's block
Arnoldi at D = 8), and the results are compared bit for bit with variant 0; then the autotuner
picks per (padded size, batch class) and the results are compared again.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rg():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import rcmps_gpu
    yield rcmps_gpu
    rcmps_gpu.gemm_variant(None)


def run_all(rg):
    from rcmps_gpu import IntegratorParameters
    from rcmps_gpu.excitation import ExcitationContext, ExcitationModel, action_batch
    out = []
    for d, seed in ((16, 11), (40, 3)):
        r, k = rg.random_state_arrays(d, scaling=0.2, seed=seed)
        s = rg.State(r, k)
        s.integrator_param.n_quad = 64
        e, G = rg.fg_phi4(s, 2.0)
        out += [np.array([e]), G]
        s.close()
    rng = np.random.default_rng(5)
    x = 0.5 * (rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8)))
    s = rg.State(0.5 * (rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8))), (x + x.conj().T) / 2,
                 steady_tol=1e-13)
    s.integrator_param = IntegratorParameters(n_quad=32, max_point=12.0, atol=1e-10, rtol=1e-10)
    ctx = ExcitationContext(s, 0.3)
    ws = [rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8)) for _ in range(3)]
    out += list(action_batch(ctx, ws, ExcitationModel("phi4", 1.0), 16, 1))
    ctx.close()
    s.close()
    return out


def test_variants_agree_bit_for_bit(rg):
    results = {}
    for v in (0, 1, 2, 3):
        note(f"case: gemm variant {v} forced")
        rg.gemm_variant(v)
        results[v] = run_all(rg)
    for v in (1, 2, 3):
        for a, b in zip(results[0], results[v]):
            assert np.array_equal(a, b), f"variant {v} differs from variant 0"
    note("case: autotuned")
    rg.gemm_variant(None)
    tuned = run_all(rg)
    for a, b in zip(results[0], tuned):
        assert np.array_equal(a, b)
    table = rg.gemm_tuning()
    for row in table:
        note(f"  Dp={row['padded_size']} class={row['batch_class']} -> variant {row['variant']} "
             f"(us: {', '.join(f'{u:.1f}' for u in row['us'])})")
        assert row["variant"] in (0, 1, 2, 3)
        assert all(u > 0 for u in row["us"])
        assert row["variant"] == 0 or row["us"][row["variant"]] < 0.9 * row["us"][0]
        assert row["variant"] == 0 or row["us"][row["variant"]] == min(row["us"])
    assert {row["padded_size"] for row in table} >= {16, 48}


def test_invalid_variant_is_rejected(rg):
    with pytest.raises(ValueError):
        rg.gemm_variant(4)
    with pytest.raises(ValueError):
        rg.gemm_variant(-2)
