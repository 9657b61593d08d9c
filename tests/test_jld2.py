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


def fields_of(jld2, path):
    d = jld2.read_jld2_datasets(path)
    integ, man = d["integration"], d["manifold"]
    return dict(R=d["R"], K=d["K"], grad=d["grad"], model=d["model"], coupling=d["coupling"], e0=d["e0"],
                info=d["info"], integration=tuple(integ[n] for n in ("n_quad", "max_point", "atol", "rtol", "maxiters")),
                manifold=tuple(man[n] for n in ("metric_reg", "precond_reg_prefactor", "precond_reg")))


@pytest.mark.parametrize("name", ["phi4_g1_D4_z2_state", "phi4_g2_D16_z2_state", "phi4_g2.7_D32_z2_state"])
def test_writer_reproduces_the_reference_checkpoints(io, tmp_path, name):
    """Rewriting a reference checkpoint from its own contents gives the same file byte for byte after
    the 512-byte preamble (whose text names the Julia version that wrote it): same objects in the same
    order, committed types and attributes, global heap and checksums."""
    io, jld2 = io
    src = FIXTURES / f"{name}.jld2"
    out = jld2.write_jld2(tmp_path / "out.jld2", fields_of(jld2, src))
    a, b = src.read_bytes(), out.read_bytes()
    assert len(a) == len(b) and a[512:] == b[512:]
    assert b[:512].startswith(b"HDF5-based Julia Data Format, version 0.2.0\x00")


def test_writer_round_trip_and_store(io, tmp_path):
    io, jld2 = io
    rng = np.random.default_rng(3)
    d = 5
    r = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    k = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    k = k + k.conj().T
    g = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    fields = dict(R=r, K=k, grad=g, model="phi4", coupling=3.25, e0=-0.125, info="round trip D=5",
                  integration=(64, 14.0, 1e-12, 1e-11, 123456), manifold=(1e-3, 1e-2, 2.5e-7))
    path = jld2.write_jld2(tmp_path / "x.jld2", fields)
    back = jld2.read_jld2_datasets(path)
    for key in ("R", "K"):
        assert np.array_equal(back[key], fields[key])
    assert np.array_equal(back["grad"], g) and back["e0"] == -0.125 and back["coupling"] == 3.25
    rec = jld2.read_jld2(path)
    assert rec.integration == io.Integration(64, 14.0, 1e-12, 1e-11, 123456)
    assert (rec.manifold.metric_reg, rec.manifold.precond_reg_prefactor, rec.manifold.precond_reg) == (1e-3, 1e-2, 2.5e-7)
    assert rec.info == "round trip D=5" and rec.energy == -0.125
    # rcmps_store on a state never evaluated: e0 NaN, zero gradient, the reference's file name
    from types import SimpleNamespace as NS
    st = NS(D=d, R=r, K=k, e0=float("nan"), grad=np.zeros((d, d), complex), info="",
            model_param=NS(name="phi4", coupling=2.0), integrator_param=NS(n_quad=32, max_point=14.0, atol=1e-13,
                                                                            rtol=1e-12, maxiters=1_000_000),
            manifold_param=NS(metric_reg=1e-3, precond_reg_prefactor=1e-2, precond_reg=1e-4))
    out = io.rcmps_store(st, directory=tmp_path / "data", id="abc")
    assert out.name == "phi4_g2.0_D5_abc.jld2"
    rec = jld2.read_jld2(out)
    assert rec.energy is None and rec.gradient is None and np.array_equal(rec.R, r)
