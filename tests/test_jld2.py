"""JLD2 checkpoints (rcmps_store schema 1, src/storage.jl:4-38) read without h5py.

Pinned by the reference's own fixtures: the D=16 checkpoint equals the portable export of the same
state byte for byte (R, K, gradient, energy); the D=4 checkpoint of a never-evaluated state has
no energy or gradient; the Z2 block structure of RandomStateZ2 (src/state.jl:61-70) holds; the
stored integration and manifold records are read with their field names; damaged files fail.
"""
import numpy as np
import pytest

from conftest import FIXTURES


@pytest.fixture(scope="module")
def io():
    import rcmps_gpu.io as io
    import rcmps_gpu.jld2 as jld2
    return io, jld2


def test_d16_checkpoint_equals_portable_export(io):
    io, jld2 = io
    a = jld2.read_jld2(FIXTURES / "phi4_g2_D16_z2_state.jld2")
    b = io.read_portable(FIXTURES / "phi4_g2_D16_z2_portable")
    assert np.array_equal(a.R, b.R) and np.array_equal(a.K, b.K) and np.array_equal(a.gradient, b.gradient)
    assert a.energy == b.energy == -0.15648780475169466
    assert (a.model, a.coupling, a.D) == ("phi4", 2.0, 16)
    assert a.integration == io.Integration(192, 20.0, 1e-13, 1e-12, 1_000_000)
    assert a.manifold.metric_reg == 1e-3 and a.manifold.precond_reg_prefactor == 1e-2


def test_unevaluated_d4_checkpoint(io):
    io, jld2 = io
    a = jld2.read_jld2(FIXTURES / "phi4_g1_D4_z2_state.jld2")
    assert a.energy is None and a.gradient is None
    assert a.integration == io.Integration(32, 14.0, 1e-13, 1e-12, 1_000_000)
    assert a.info.startswith("RandomStateZ2(4; seed=11, scaling=0.5)")
    # Z2: R odd, K even under diag(1, 1, -1, -1)
    assert np.all(a.R[:2, :2] == 0) and np.all(a.R[2:, 2:] == 0)
    assert np.all(a.K[:2, 2:] == 0) and np.all(a.K[2:, :2] == 0)
    np.testing.assert_allclose(a.K, a.K.conj().T, atol=1e-15)


def test_d32_checkpoint(io):
    io, jld2 = io
    a = jld2.read_jld2(FIXTURES / "phi4_g2.7_D32_z2_state.jld2")
    assert a.D == 32 and a.coupling == 2.7 and a.energy == -0.30190930265518967
    assert a.gradient.shape == (32, 32) and np.abs(a.gradient).max() < 1e-4
    np.testing.assert_allclose(a.K, a.K.conj().T, atol=1e-14)


def test_damaged_files(io, tmp_path):
    io, jld2 = io
    data = (FIXTURES / "phi4_g2_D16_z2_state.jld2").read_bytes()
    bad = tmp_path / "bad.jld2"
    bad.write_bytes(data[:400])
    with pytest.raises(jld2.JLD2Error):
        jld2.read_jld2(bad)
    bad.write_bytes(data[:700] + bytes(len(data) - 700))
    with pytest.raises(jld2.JLD2Error):
        jld2.read_jld2(bad)
