"""fg_phi4 energy+gradient evaluations per second on one MI355X.

Workload: the D=64 phi^4 state of the headline metric: R, K complex normal scaled 0.2 (K
Hermitised, seed 42, numpy PCG64 draws in RandomState order), g = 2.0, n_quad = 256 (512 Gauss
nodes), max_point = 20, atol = 1e-11, rtol = 1e-10.  A step is one full fg_phi4 evaluation
(six DOP853 trajectories, pullback, boundary Poisson solve, metric solve) on a prepared state
whose inputs are resident in HBM.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (str(ROOT), str(ROOT / "rcmpscore.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

FP64_MFMA_PEAK_TFLOPS = 78.6
METRIC = "fg_phi4 E+grad evals/s at D=64, split-prec vs FP64 on GPU at |ΔE|/|E|≤1e-10"
# complex D x D products per right-hand side of each reduced system (level-1 and level-2 tiles):
# kinetic fwd, kinetic adj, kinetic-commutator fwd, kinetic-commutator adj, phi4 fwd, phi4 adj
GEMMS_PER_RHS = (4, 8, 4, 9, 12, 16)


def workload(D=64, g=2.0, seed=42, scaling=0.2, n_quad=256, max_point=20.0, atol=1e-11, rtol=1e-10):
    return dict(D=D, g=g, seed=seed, scaling=scaling, n_quad=n_quad, max_point=max_point, atol=atol, rtol=rtol)


def profiled_traffic(wl):
    """HBM bytes per trajectory launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of
    this workload (copy-calibrated), or None when no summary matches the workload."""
    if (wl["D"], wl["n_quad"]) != (64, 256):
        return None
    path = ROOT / "profiles" / "rc_trajectory_D64_counters.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(R, Q, rho, wl, gpu_stats, attempts_sample=10):
    """Oracle (FP64 C port) timed on a bounded sample: `attempts_sample` DOP853 attempts of each
    of the six trajectories, run concurrently on six host threads; the full evaluation time is
    extrapolated with the attempt/dense-step counts of the GPU run, whose step sequences equal the
    oracle's.  Pullback and boundary solve are not included (a lower bound on the oracle's time)."""
    import oracle

    def one(system):
        t0 = time.perf_counter()
        try:
            oracle.trajectory(R, Q, rho, system, n_quad=wl["n_quad"], max_point=wl["max_point"],
                              atol=wl["atol"], rtol=wl["rtol"], max_steps=attempts_sample)
        except oracle.OracleFailure:
            pass
        return time.perf_counter() - t0

    t0 = time.perf_counter()
    with ThreadPoolExecutor(6) as ex:
        times = list(ex.map(one, range(6)))
    wall = time.perf_counter() - t0
    est = 0.0
    for s in range(6):
        per_rhs = times[s] / (1 + 12 * attempts_sample)
        rhs = 1 + 12 * gpu_stats["attempts"][s] + 3 * gpu_stats["dense_steps"][s]
        est = max(est, per_rhs * rhs)
    return dict(value=1.0 / est, unit="evals/s", cores=6, kind="port",
                sample=(f"{attempts_sample} DOP853 attempts of each of the 6 trajectories of the D=64 "
                        f"workload on 6 threads ({wall:.1f} s wall, {sum(times):.1f} s CPU); evaluation "
                        f"time extrapolated from the GPU run's attempt counts, pullback/boundary excluded"),
                estimated_seconds_per_eval=est)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--D", type=int, default=64)
    ap.add_argument("--n-quad", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-attempts", type=int, default=10)
    ap.add_argument("--copy-reference", type=int, default=0, metavar="MIB",
                    help="first run one plain device copy of this many MiB (a known-size kernel that "
                         "calibrates HBM byte counters in a profile of the same command)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(independent states shard across GPUs with no exchange)", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    import rcmps_gpu

    if args.copy_reference > 0:
        n = args.copy_reference * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    wl = workload(D=args.D, n_quad=args.n_quad)
    r, k = rcmps_gpu.random_state_arrays(wl["D"], scaling=wl["scaling"], seed=wl["seed"])
    t0 = time.perf_counter()
    s = rcmps_gpu.State(r, k)
    prep_s = time.perf_counter() - t0
    ip = s.integrator_param
    ip.n_quad, ip.max_point, ip.atol, ip.rtol = wl["n_quad"], wl["max_point"], wl["atol"], wl["rtol"]
    for _ in range(args.warmup):
        rcmps_gpu.fg_phi4_raw(s, wl["g"])
    torch.cuda.synchronize()
    traj_ms, energies, stats = [], [], None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e, G, _, _, stats = rcmps_gpu.fg_phi4_raw(s, wl["g"])
        traj_ms.append(stats["trajectory_ms"])
        energies.append(e)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = args.steps / elapsed
    D = wl["D"]
    flops = sum(stats["rhs_calls"][i] * GEMMS_PER_RHS[i] for i in range(6)) * 8.0 * D**3
    avg_traj_s = sum(traj_ms) / len(traj_ms) / 1e3
    achieved = flops / avg_traj_s / 1e12
    roofline = dict(bound="mfma", achieved=achieved, peak=FP64_MFMA_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=achieved / FP64_MFMA_PEAK_TFLOPS, traffic=profiled_traffic(wl),
                    kernel="rc_trajectory<4>", kernel_ms=avg_traj_s * 1e3,
                    algorithmic_tflop_per_launch=flops / 1e12)
    out = {
        "metric": METRIC, "value": value, "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"fg_phi4 D={D} g={wl['g']} seeded random state (scaling {wl['scaling']}, "
                               f"seed {wl['seed']}), n_quad={wl['n_quad']}, max_point={wl['max_point']}, "
                               f"atol={wl['atol']}, rtol={wl['rtol']}",
                   "precision": "fp64", "D": D, "n_quad": wl["n_quad"], "g": wl["g"]},
        "roofline": roofline,
        "energy": energies[-1],
        "energy_spread": max(energies) - min(energies),
        "attempts": stats["attempts"], "dense_steps": stats["dense_steps"],
        "trajectory_ms": avg_traj_s * 1e3, "pullback_ms": stats["pullback_ms"],
        "boundary_ms": stats["boundary_ms"], "state_prep_s": prep_s,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(s.R, s.Q, s.rho, wl, stats, args.cpu_attempts)
    print(json.dumps(out), flush=True)
    s.close()
    torch.cuda.synchronize()
    return 0


if __name__ == "__main__":
    rc = main()
    if rc:
        sys.exit(rc)
