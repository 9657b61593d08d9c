"""rcmps_boundary_environment_noalloc, the boundary Poisson solve on the library's fixed device array,
against rcmps_boundary_environment (same kernels, same order: the same bits) through the C ABI.

Two no-alloc calls run back to back, the second on a smaller state, so the second one meets the
first one's leftovers in the shared array.  A state beyond the array (D = 65, padded 80) is refused.
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


def hermitian(d, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    return a + a.conj().T


def test_noalloc_boundary_equals_the_allocating_entry_point(rg):
    cases = []
    for d, seed in ((64, 21), (12, 22)):
        note(f"case: no-alloc boundary reference D={d} seed={seed}")
        r, k = rg.random_state_arrays(d, scaling=0.2, seed=seed)
        s = rg.State(r, k)
        a = hermitian(d, seed + 100)
        cases.append((s, a, rg.boundary_environment(s, a, tol=1e-12)))
    # the no-alloc calls back to back: D = 64 (3.2 MiB of the array), then D = 12 on its leftovers
    for s, a, h_ref in cases:
        note(f"case: no-alloc boundary D={s.D}")
        h = rg.boundary_environment(s, a, tol=1e-12, no_alloc=True)
        assert h.shape == h_ref.shape
        assert np.array_equal(h.view(np.float64), h_ref.view(np.float64)), np.abs(h - h_ref).max()
    # and the small solve against the oracle's dense solve
    s, a, _ = cases[1]
    h = rg.boundary_environment(s, a, tol=1e-12, no_alloc=True)
    ref = oracle.boundary_environment(s.Q, s.R, s.rho, a, method="dense")
    np.testing.assert_allclose(h, ref, atol=1e-9 * max(1.0, np.abs(ref).max()))
    for s, _, _ in cases:
        s.close()


def test_noalloc_boundary_refuses_a_state_beyond_its_scratch(rg):
    note("case: no-alloc boundary D=65 refused")
    r, k = rg.random_state_arrays(65, scaling=0.2, seed=1)
    s = rg.State(r, k)
    with pytest.raises(ValueError, match="scratch"):
        rg.boundary_environment(s, hermitian(65, 3), no_alloc=True)
    s.close()
