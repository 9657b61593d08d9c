"""The measured fg_phi4 workloads on the GPU against the oracle, through the C ABI.

- The throughput state (D = 64, g = 2, seed 42, scaling 0.2, n_quad 256, window 20, 1e-11/1e-10)
  and the D = 32 state of the same distribution at n_quad 256: both precisions against the oracle's
  FP64 result committed in tests/golden/fg_phi4_references.npz (written by
  tests/golden/make_fg_phi4_references.py).  The GPU state takes the oracle's rho0, so both sides
  integrate the same state.  FP64: energy within 1e-12 relative, gradient within 1e-10 of max |G|,
  and the oracle's accepted / rejected / attempt counts per system.  Split precision: energy and
  gradient within 1e-10.
- The 256-state batch (D uniform in {8, 16, 24, 32, 48, 64}, g uniform in [0.5, 4], seeds 0..255,
  default integrator), one call per precision: split against FP64 within 1e-10, one single-state
  call per D equal to the batch bit for bit in both precisions, and the D = 8 members against the
  oracle (1e-12 / 1e-10).
- The reference's D = 16 portable fixture at tightened tolerances (1e-15/1e-14): the FP64 GPU energy
  equals its e0 = -0.15648780475169466 to 1e-12 relative.
"""
import sys

import numpy as np
import pytest

import oracle
from conftest import FIXTURES

pytestmark = pytest.mark.gpu

REFS = FIXTURES / "fg_phi4_references.npz"
CASES = {"bench_D64": dict(D=64, seed=42, scaling=0.2, g=2.0, n_quad=256),
         "cfg2_D32": dict(D=32, seed=42, scaling=0.2, g=2.0, n_quad=256)}


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rg():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import rcmps_gpu
    return rcmps_gpu


def rel_grad(g, ref):
    return np.abs(g - ref).max() / np.abs(ref).max()


@pytest.mark.parametrize("name", sorted(CASES))
def test_workload_state_matches_the_oracle(rg, name):
    c = CASES[name]
    note(f"case: {name} D={c['D']} n_quad={c['n_quad']} fp64 and split against the committed oracle result")
    with np.load(REFS) as z:
        ref = {key.split("/", 1)[1]: z[key] for key in z.files if key.startswith(name + "/")}
    r, k = rg.random_state_arrays(c["D"], scaling=c["scaling"], seed=c["seed"])
    np.testing.assert_allclose([r.sum(), k.sum()], ref["rk_sums"], rtol=0, atol=1e-12)
    s = rg.State(r, k, rho=ref["rho"])
    s.integrator_param.n_quad = c["n_quad"]
    assert (s.integrator_param.max_point, s.integrator_param.atol, s.integrator_param.rtol) == (20.0, 1e-11, 1e-10)
    e_ref, g_ref = float(ref["energy"]), ref["gradient"]
    e, g, st = rg.fg_phi4(s, c["g"], return_stats=True)
    note(f"  fp64: dE/E = {abs(e - e_ref) / abs(e_ref):.2e}, dG = {rel_grad(g, g_ref):.2e}")
    assert abs(e - e_ref) <= 1e-12 * abs(e_ref)
    assert rel_grad(g, g_ref) <= 1e-10
    counts = ref["stats"]   # per system: accepted, rejected, attempts, dense steps, rhs calls
    assert list(st["accepted"]) == counts[:, 0].tolist()
    assert list(st["rejected"]) == counts[:, 1].tolist()
    assert list(st["attempts"]) == counts[:, 2].tolist()
    es, gs, sts = rg.fg_phi4(s, c["g"], return_stats=True, precision="split")
    note(f"  split: dE/E = {abs(es - e_ref) / abs(e_ref):.2e}, dG = {rel_grad(gs, g_ref):.2e}, "
         f"attempts {list(sts['attempts'])}")
    assert abs(es - e_ref) <= 1e-10 * abs(e_ref)
    assert rel_grad(gs, g_ref) <= 1e-10
    s.close()


def batch_workload(rg, n=256):
    rng = np.random.default_rng(0)
    dims = rng.choice([8, 16, 24, 32, 48, 64], size=n)
    gs = [float(x) for x in rng.uniform(0.5, 4.0, size=n)]
    states = []
    for i, d in enumerate(dims):
        r, k = rg.random_state_arrays(int(d), scaling=0.2, seed=i)
        states.append(rg.State(r, k))
    return states, gs


def test_256_state_batch_both_precisions(rg):
    note("case: 256-state batch, D in {8..64}, g in [0.5, 4], seeds 0..255, fp64 and split")
    states, gs = batch_workload(rg)
    out = {}
    for precision in ("fp64", "split"):
        E, G, stats = rg.fg_phi4_batch(states, gs, precision=precision, return_stats=True)
        assert all(x["status"] == 0 for x in stats)
        out[precision] = (np.array(E), G)
    (Ef, Gf), (Es, Gs) = out["fp64"], out["split"]
    e_err = np.max(np.abs(Es - Ef) / np.abs(Ef))
    g_err = max(rel_grad(a, b) for a, b in zip(Gs, Gf))
    note(f"  split vs fp64: energy {e_err:.2e}, gradient {g_err:.2e}")
    assert e_err <= 1e-10 and g_err <= 1e-10
    seen = set()
    for i, s in enumerate(states):
        if s.D in seen:
            continue
        seen.add(s.D)
        for precision in ("fp64", "split"):
            e1, g1 = rg.fg_phi4(s, gs[i], precision=precision)
            E, G = out[precision]
            assert E[i] == e1 and np.array_equal(G[i], g1), (i, s.D, precision)
    assert seen == {8, 16, 24, 32, 48, 64}
    checked = 0
    for i, s in enumerate(states):
        if s.D != 8 or checked == 6:
            continue
        ref = oracle.fg_phi4(s.R, s.Q, s.rho, gs[i])
        assert Ef[i] == pytest.approx(ref["energy"], rel=1e-12)
        assert rel_grad(Gf[i], ref["gradient"]) <= 1e-10
        checked += 1
    assert checked == 6
    for s in states:
        s.close()


def test_portable_fixture_energy_at_tight_tolerances(rg):
    note("case: D16 portable fixture, fp64 at 1e-15/1e-14, n_quad 192")
    d = FIXTURES / "phi4_g2_D16_z2_portable"
    ld = lambda name: np.fromfile(d / f"{name}.bin", dtype="<c16").reshape((16, 16), order="F")  # noqa: E731
    s = rg.State(ld("R"), ld("K"), rho=ld("rho"))
    s.integrator_param.n_quad = 192
    s.integrator_param.atol = 1e-15
    s.integrator_param.rtol = 1e-14
    e, _ = rg.fg_phi4(s, 2.0)
    note(f"  E = {e:.17g}, relative to e0 {abs(e + 0.15648780475169466) / 0.15648780475169466:.2e}")
    assert e == pytest.approx(-0.15648780475169466, rel=1e-12)
    s.close()
