"""Portable state format 1 (reference spec gpu/src/rcmps_gpu/io.py:1-104): reading the shared
fixture, byte-exact round trips, and the format's error cases."""
import math
import shutil

import numpy as np
import pytest

from conftest import FIXTURES
from rcmps_gpu import io

FIX = FIXTURES / "phi4_g2_D16_z2_portable"


def test_reads_the_shared_fixture():
    rec = io.read_portable(FIX)
    assert (rec.D, rec.model, rec.coupling, rec.mass) == (16, "phi4", 2.0, 1.0)
    assert rec.energy == -0.15648780475169466
    assert rec.integration == io.Integration(192, 20.0, 1e-13, 1e-12, 1000000)
    assert rec.manifold.metric_reg == 1e-3 and rec.manifold.precond_reg_prefactor == 0.01
    raw = np.fromfile(FIX / "R.bin", dtype="<c16")
    assert rec.R[3, 0] == raw[3] and rec.R[0, 3] == raw[3 * 16]       # column-major on disk
    # the stored gradient norm is sqrt(metric(G, G)) with metric_reg of [manifold]
    G, rho = rec.gradient, rec.rho
    gn = math.sqrt(2 * np.vdot(G, G @ (rho + rec.manifold.metric_reg * np.eye(16))).real)
    assert gn == pytest.approx(rec.gradient_norm, rel=1e-12)
    # K is Hermitian, rho is a density
    np.testing.assert_allclose(rec.K, rec.K.conj().T, atol=1e-14)
    assert np.trace(rho) == pytest.approx(1.0, abs=1e-12)


def test_round_trip_is_byte_exact(tmp_path):
    rec = io.read_portable(FIX)
    out = io.write_portable(tmp_path / "copy", rec)
    for name in ("R", "K", "rho", "gradient"):
        assert (out / f"{name}.bin").read_bytes() == (FIX / f"{name}.bin").read_bytes()
    back = io.read_portable(out)
    for field in ("D", "model", "coupling", "mass", "energy", "gradient_norm", "integration", "manifold",
                  "producer", "source", "info"):
        assert getattr(back, field) == getattr(rec, field), field
    text = (out / "metadata.toml").read_text()
    assert text.index("format =") < text.index("D =") < text.index("model =") < text.index("energy =")


def test_format_errors(tmp_path):
    bad = tmp_path / "bad"
    shutil.copytree(FIX, bad)
    (bad / "R.bin").write_bytes(b"\0" * 100)
    with pytest.raises(ValueError, match="bytes"):
        io.read_portable(bad)
    shutil.rmtree(bad)
    shutil.copytree(FIX, bad)
    (bad / "K.bin").unlink()
    with pytest.raises(ValueError, match="K.bin"):
        io.read_portable(bad)
    shutil.rmtree(bad)
    shutil.copytree(FIX, bad)
    text = (bad / "metadata.toml").read_text().replace('format = "rcmps-portable-state"', 'format = "other"')
    (bad / "metadata.toml").write_text(text)
    with pytest.raises(ValueError, match="format"):
        io.read_portable(bad)
    rec = io.read_portable(FIX)
    with pytest.raises(FileExistsError):
        io.write_portable(FIX, rec)
    odd = io.StateRecord(D=2, model="phi4", coupling=1.0, R=np.eye(3), K=np.eye(2))
    with pytest.raises(ValueError):
        io.write_portable(tmp_path / "odd", odd)
