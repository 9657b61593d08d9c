"""Split precision (INT8 MFMA over exact 8-bit slices of the FP64 operands) against the FP64
path and the oracle, through the C ABI.

Bars: energy within 1e-10 relative of the FP64 GPU path and of the oracle (the accuracy the
split-precision throughput is quoted at); gradient within 1e-8 relative (to max |G|) of both;
both errors are printed.  The slices carry 48 bits per operand and exact INT32 sums, so the
products are accurate to about 2^-45 relative to max|A| max|B| k; the difference from FP64 is
dominated by the adaptive step sequence, which must stay within a few attempts of FP64's.
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


@pytest.mark.parametrize("d,seed,scaling,n_quad", [(4, 7, 0.3, 32), (8, 42, 0.2, 128), (16, 11, 0.2, 64),
                                                    (24, 5, 0.2, 64), (40, 3, 0.2, 64)])
def test_split_matches_fp64_and_oracle(rg, d, seed, scaling, n_quad):
    note(f"case: split D={d} seed={seed} n_quad={n_quad}")
    r, k = rg.random_state_arrays(d, scaling=scaling, seed=seed)
    s = rg.State(r, k)
    s.integrator_param.n_quad = n_quad
    e64, G64, st64 = rg.fg_phi4(s, 2.0, return_stats=True)
    esp, Gsp, stsp = rg.fg_phi4(s, 2.0, return_stats=True, precision="split")
    de = abs(esp - e64) / abs(e64)
    dg = np.abs(Gsp - G64).max() / np.abs(G64).max()
    note(f"  |dE|/|E| = {de:.2e}  |dG|/max|G| = {dg:.2e}  attempts fp64={st64['attempts']} "
         f"split={stsp['attempts']}")
    assert de <= 1e-10
    assert dg <= 1e-8
    if d <= 16:
        ref = oracle.fg_phi4(s.R, s.Q, s.rho, 2.0, n_quad=n_quad)
        de_o = abs(esp - ref["energy"]) / abs(ref["energy"])
        dg_o = np.abs(Gsp - ref["gradient"]).max() / np.abs(ref["gradient"]).max()
        note(f"  vs oracle: |dE|/|E| = {de_o:.2e}  |dG|/max|G| = {dg_o:.2e}")
        assert de_o <= 1e-10
        assert dg_o <= 1e-8
    assert all(abs(a - b) <= 2 + 0.02 * a for a, b in zip(st64["attempts"], stsp["attempts"]))
    s.close()


def test_split_d64_workload(rg):
    note("case: split D=64 g=2 seed=42 n_quad=256")
    r, k = rg.random_state_arrays(64, scaling=0.2, seed=42)
    s = rg.State(r, k)
    s.integrator_param.n_quad = 256
    e64, G64, st64 = rg.fg_phi4(s, 2.0, return_stats=True)
    esp, Gsp, stsp = rg.fg_phi4(s, 2.0, return_stats=True, precision="split")
    de = abs(esp - e64) / abs(e64)
    dg = np.abs(Gsp - G64).max() / np.abs(G64).max()
    note(f"  E64={e64!r} Esplit={esp!r} |dE|/|E| = {de:.2e}  |dG|/max|G| = {dg:.2e}")
    note(f"  trajectory ms fp64={st64['trajectory_ms']:.2f} split={stsp['trajectory_ms']:.2f}")
    assert de <= 1e-10
    assert dg <= 1e-8
    s.close()
