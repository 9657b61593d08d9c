"""Writes fg_phi4_references.npz: the oracle's FP64 fg_phi4 on the seeded states of the measured
workloads, for the GPU parity tests that cannot run the oracle in the time a test may take.

States (numpy PCG64 draws in RandomState order, src/state.jl:55-60, as rcmps_gpu.random_state_arrays):
  bench_D64: D = 64, scaling 0.2, seed 42, g = 2, n_quad 256, window 20, atol 1e-11, rtol 1e-10
             (the fg_phi4 throughput workload)
  cfg2_D32:  D = 32, scaling 0.2, seed 42, g = 2, n_quad 256, same window and tolerances

For each: rho0 from the oracle's CPTP Arnoldi at tol 1e-12 (the GPU tests pass it to State(...,
rho=...), so both sides integrate the same state), the energy, the Riemannian gradient (metric_reg
1e-3), the term values, and per system the oracle's (accepted, rejected, attempts, dense steps, rhs calls).
R and K are regenerated from the seed by the tests; their sums are stored to catch a changed generator.

Run from the repository root: python tests/golden/make_fg_phi4_references.py
"""
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT), str(ROOT / "rcmpscore.jl_amd")]

import oracle  # noqa: E402
from rcmps_gpu.state import random_state_arrays  # noqa: E402

CASES = {
    "bench_D64": dict(D=64, seed=42, scaling=0.2, g=2.0, n_quad=256),
    "cfg2_D32": dict(D=32, seed=42, scaling=0.2, g=2.0, n_quad=256),
}


def main():
    out = {}
    for name, c in CASES.items():
        t0 = time.time()
        r, k = random_state_arrays(c["D"], scaling=c["scaling"], seed=c["seed"])
        q = oracle.canonical_q(r, k)
        rho = oracle.steady_state(q, r, tol=1e-12)
        res = oracle.fg_phi4(r, q, rho, c["g"], n_quad=c["n_quad"], max_point=20.0, atol=1e-11, rtol=1e-10,
                             nthreads=8)
        out[f"{name}/rho"] = rho
        out[f"{name}/energy"] = np.float64(res["energy"])
        out[f"{name}/gradient"] = res["gradient"]
        out[f"{name}/term_values"] = res["term_values"]
        out[f"{name}/stats"] = res["stats"]
        out[f"{name}/rk_sums"] = np.array([r.sum(), k.sum()])
        print(f"{name}: E = {res['energy']:.17g}, attempts {res['stats'][:, 2].tolist()}, "
              f"{time.time() - t0:.1f} s", flush=True)
    path = Path(__file__).with_name("fg_phi4_references.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
