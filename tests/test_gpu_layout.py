"""Column-major matrices at the C ABI (rcmps_set_layout), as a Julia caller holds them.

A context set to RCMPS_LAYOUT_COLUMN_MAJOR takes and returns every D x D matrix in Julia's layout;
the device works in padded row-major either way, so the results are the row-major context's bits,
transposed: stationary state, fg_phi4 energy, gradient, covector and environment, the boundary
solve, and a two-state batch, at D = 5 (one tile) and D = 40 (three tiles per side, padded).
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import rcmps_gpu
    from rcmps_gpu import _lib
    lib = _lib.load()
    ctx = ctypes.c_void_p()
    _lib.check(lib.rcmps_context_create(0, ctypes.byref(ctx)), "context")
    _lib.check(lib.rcmps_set_layout(ctx, 1), "layout")
    yield rcmps_gpu, _lib, lib, ctx, torch
    lib.rcmps_context_destroy(ctx)


def colmajor(torch, a):
    """device buffer holding `a` column-major"""
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.complex128).T)).to("cuda")


def from_colmajor(t, d):
    return t.cpu().numpy().reshape(d, d).T


@pytest.mark.parametrize("d", [5, 40])
def test_column_major_equals_row_major(env, d):
    rg, _lib, lib, ctx, torch = env
    note(f"case: column-major layout D={d}")
    r, k = rg.random_state_arrays(d, scaling=0.3 if d == 5 else 0.2, seed=7)
    s = rg.State(r, k, steady_tol=1e-12)
    e_row, G_row, cov_row, env_row, _ = rg.fg_phi4_raw(s, 2.0)
    h_row = rg.boundary_environment(s, s.Q + s.Q.conj().T, tol=1e-12)

    R, K = colmajor(torch, r), colmajor(torch, k)
    st = ctypes.c_void_p()
    stats = _lib.Stats()
    torch.cuda.synchronize()
    _lib.check(lib.rcmps_state_create(ctx, d, R.data_ptr(), K.data_ptr(), None, 1e-12, ctypes.byref(st),
                                      ctypes.byref(stats)), "state", stats)
    try:
        rho = torch.empty(d * d, dtype=torch.complex128, device="cuda")
        _lib.check(lib.rcmps_state_get(ctx, st, 1, rho.data_ptr()), "rho")
        torch.cuda.synchronize()
        assert np.array_equal(from_colmajor(rho, d), s.rho)
        ip = _lib.Integrator(s.integrator_param.n_quad, s.integrator_param.max_point, s.integrator_param.atol,
                             s.integrator_param.rtol, s.integrator_param.maxiters, 0.01)
        G, cov, envm = (torch.empty(d * d, dtype=torch.complex128, device="cuda") for _ in range(3))
        e = ctypes.c_double()
        _lib.check(lib.rcmps_fg_phi4(ctx, st, 2.0, ctypes.byref(ip), s.manifold_param.metric_reg, 0, ctypes.byref(e),
                                     G.data_ptr(), cov.data_ptr(), envm.data_ptr(), ctypes.byref(stats)), "fg", stats)
        torch.cuda.synchronize()
        note(f"  E = {e.value!r} (row-major context {e_row!r})")
        assert e.value == e_row
        assert np.array_equal(from_colmajor(G, d), G_row.cpu().numpy())
        assert np.array_equal(from_colmajor(cov, d), cov_row.cpu().numpy())
        assert np.array_equal(from_colmajor(envm, d), env_row.cpu().numpy())
        a = colmajor(torch, s.Q + s.Q.conj().T)
        h = torch.empty(d * d, dtype=torch.complex128, device="cuda")
        _lib.check(lib.rcmps_boundary_environment(ctx, st, a.data_ptr(), 1e-12, h.data_ptr(), ctypes.byref(stats)),
                   "boundary", stats)
        torch.cuda.synchronize()
        assert np.array_equal(from_colmajor(h, d), h_row)
    finally:
        lib.rcmps_state_destroy(st)
    s.close()


def test_column_major_batch(env):
    rg, _lib, lib, ctx, torch = env
    note("case: column-major layout, two-state batch")
    dims, gs = (8, 24), (1.5, 3.0)
    rows, handles, keep = [], [], []
    for i, d in enumerate(dims):
        r, k = rg.random_state_arrays(d, scaling=0.2, seed=20 + i)
        rows.append(rg.State(r, k))
        R, K = colmajor(torch, r), colmajor(torch, k)
        keep += [R, K]
        st = ctypes.c_void_p()
        stats = _lib.Stats()
        torch.cuda.synchronize()
        _lib.check(lib.rcmps_state_create(ctx, d, R.data_ptr(), K.data_ptr(), None, 1e-11, ctypes.byref(st),
                                          ctypes.byref(stats)), "state", stats)
        handles.append(st)
    try:
        e_row, g_row = rg.fg_phi4_batch(rows, gs)
        ip = rg.observables._integrator(rows[0].integrator_param)
        grads = [torch.empty(d * d, dtype=torch.complex128, device="cuda") for d in dims]
        energies = (ctypes.c_double * 2)()
        stats = (_lib.Stats * 2)()
        _lib.check(lib.rcmps_fg_phi4_batch(ctx, 2, (ctypes.c_void_p * 2)(*[h.value for h in handles]),
                                           (ctypes.c_double * 2)(*gs), ctypes.byref(ip),
                                           rows[0].manifold_param.metric_reg, 0, energies,
                                           (ctypes.c_void_p * 2)(*[g.data_ptr() for g in grads]), stats), "batch")
        torch.cuda.synchronize()
        for i, d in enumerate(dims):
            assert energies[i] == e_row[i]
            assert np.array_equal(from_colmajor(grads[i], d), g_row[i])
    finally:
        for h in handles:
            lib.rcmps_state_destroy(h)
        for s in rows:
            s.close()
