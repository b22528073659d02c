#!/usr/bin/env python3
"""Throughput of the plain FP32 Bloch-Torrey solve on one MI355X.

Workload: 256^3 voxels (1 um), 20% pores from the seeded Gaussian-smoothed
porosity generator, divergence-free pore velocity (max 1 mm/s), pulsed
gradient along z (G = 0.3 T/m, delta 2 ms, Delta 20 ms), complex (Mx, My)
in FP32, D = 2.3e-9 m^2/s, T2 = 2 s, rho = 10 um/s, dt = 0.4 x the explicit
limit.  A step is one time step of the whole grid (update + per-step signal
reduction); inputs are resident in HBM before the timed region.

Prints one JSON line.  The roofline object prices the sweep kernel at its
compulsory traffic, 29 bytes per voxel-step (read Mx, My: 8; write Mx, My:
8; uint8 mask: 1; Vx, Vy, Vz faces: 12), against 8 TB/s.  cpu_baseline
times the FP64 CPU oracle on this host on a bounded number of steps of the
same grid.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "voxel-steps/s & %HBM roofline @256³; learned vs plain time-to-signal @1e-3"
HBM_PEAK_GBS = 8000.0
BYTES_PER_VOXEL_STEP = 29.0  # read Mx,My (8) + write Mx,My (8) + mask (1) + Vx,Vy,Vz (12)


def cpu_baseline(case, target_s: float):
    """FP64 oracle steps of the same grid on the host cores (bounded sample)."""
    import numpy as np

    import oracle

    p, act, D, t1, t2, vel, m = oracle.setup_case(case, m0=[case.mask.astype(np.float64),
                                                           np.zeros(case.shape), np.zeros(case.shape)])
    t0 = time.perf_counter()
    oracle.run(p, act, D, t1, t2, vel, 1, m, grad_sched=case.grad_sched[:1])
    one = time.perf_counter() - t0
    n = max(1, min(int(target_s / max(one, 1e-6)), 200))
    t0 = time.perf_counter()
    oracle.run(p, act, D, t1, t2, vel, n, m, grad_sched=case.grad_sched[1 : n + 1])
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": case.nvox * n / dt,
        "unit": "voxel-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"FP64 oracle (OpenMP restatement of ExplicitSimple), {case.shape[2]}^3, {n} steps, "
                  f"{dt:.1f} s, {threads} threads",
    }


def copy_probe(mib: int) -> int:
    """Plain copy of a known size: its GB/s is the practical HBM ceiling, and
    its FETCH_SIZE / WRITE_SIZE counters give the bytes-per-KiB scale."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 10
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"copy_mib": mib, "ms": ms, "GBps": 2 * n * 4 / (ms * 1e-3) / 1e9}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=256, help="grid edge")
    ap.add_argument("--phi", type=float, default=0.20)
    ap.add_argument("--variant", type=int, default=-1)
    ap.add_argument("--lz", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="time a plain device-to-device copy of MIB MiB instead (HBM ceiling / counter scale)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU; the solve keeps the whole domain on one MI355X "
              "(multi-GPU z-slab sharding is a design note, not built)", file=sys.stderr)
        return 2

    from initializator_amd import generators
    from initializator_amd.solver import HipSolver

    total = args.warmup + args.steps
    case = generators.case_flow(n=args.n, phi=args.phi, nsteps=total)
    s = HipSolver.from_case(case)
    s.set_tile(args.variant, args.lz)
    s.upload_magnetization(case.mask.astype("float64"), np.zeros(case.shape))
    s.run(args.warmup, case.grad_sched[: args.warmup], want_signal=False)
    t0 = time.perf_counter()
    s.run(args.steps, case.grad_sched[args.warmup :], want_signal=True)
    wall = time.perf_counter() - t0
    gpu_ms, nsw = s.last_run_ms()
    s.close()

    nvox = case.nvox
    value = nvox * args.steps / wall
    kernel_s = gpu_ms * 1e-3 / args.steps
    achieved = BYTES_PER_VOXEL_STEP * nvox / kernel_s / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded porosity, divergence-free velocity, PGSE schedule)",
        "config": {
            "workload": f"plain FP32 upwind Bloch-Torrey step, {args.n}^3, phi={args.phi}, flow + PGSE "
                        "G=0.3 T/m, complex (Mx,My)",
            "grid": [args.n] * 3,
            "bytes_per_voxel_step": BYTES_PER_VOXEL_STEP,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel_us": kernel_s * 1e6,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        cb = generators.case_flow(n=args.n, phi=args.phi, nsteps=201)
        out["cpu_baseline"] = cpu_baseline(cb, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
