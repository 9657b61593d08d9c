"""bench.py's command-line contract on the CPU (no GPU is touched).

Checks: more than one GPU is refused with a message and exit status 2; negative step counts are
refused; the default workload is the D=64 state (seed 42, scaling 0.2, g = 2, 256 nodes, window
20, 1e-11/1e-10); the roofline traffic is read from the committed copy-calibrated counter summary
of that workload, and is absent for any other workload.
"""
import importlib.util
import pathlib

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused(bench, capsys):
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU" in capsys.readouterr().err
    assert bench.main(["--gpus", "8", "--steps", "3"]) == 2


def test_invalid_step_counts_are_refused(bench, capsys):
    assert bench.main(["--steps", "0"]) == 2
    assert bench.main(["--warmup", "-1"]) == 2
    assert "--steps" in capsys.readouterr().err


def test_default_workload_and_profiled_traffic(bench):
    wl = bench.workload()
    assert (wl["D"], wl["n_quad"], wl["seed"], wl["scaling"], wl["g"]) == (64, 256, 42, 0.2, 2.0)
    assert (wl["max_point"], wl["atol"], wl["rtol"]) == (20.0, 1e-11, 1e-10)
    traffic = bench.profiled_traffic(wl)
    # copy-calibrated FETCH_SIZE + WRITE_SIZE of the split trajectory kernel: tens of GB per launch,
    # far above the 0.44 GB of dense-output samples it must write
    assert traffic is not None and 1e10 < traffic < 1e11
    assert bench.profiled_traffic(bench.workload(D=32)) is None
    assert bench.profiled_traffic(bench.workload(n_quad=128)) is None
