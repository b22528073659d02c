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
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "voxel-steps/s & %HBM roofline @256³; learned vs plain time-to-signal @1e-3"
HBM_PEAK_GBS = 8000.0
BYTES_PER_VOXEL_STEP = 29.0  # read Mx,My (8) + write Mx,My (8) + mask (1) + velocity (12)
# MacCormack (--scheme maccormack): predictor pass 29 B (as above, writing the
# predictor) + corrector pass 37 B (reads the predictor and the state, 8 each)
BYTES_PER_VOXEL_STEP_MAK = 66.0
# Bytes past the L2 per sweep launch of the default workload (256^3, phi
# 0.20, automatic tile, brick layout): rocprofv3 --pmc FETCH_SIZE (x2.0) +
# WRITE_SIZE (x1.0) per dispatch, scaled by what the same counters read for a
# plain 1 GiB copy (summary under profiles/, sweep_kernel_256cube_flow_pgse.md).
# The device's memory-side cache serves part of these requests; the counters
# at the L2 do not tell which.
DEFAULT_TRAFFIC = 226.1e6 + 52.5e6
# The same for --scheme maccormack at 256^3, phi 0.20 (brick layout): the
# predictor and the corrector launch of one step together (summary under
# profiles/, gsweep_maccormack_256cube_flow_pgse.md).
MAK_TRAFFIC = (219.2e6 + 50.4e6) + (286.7e6 + 50.4e6)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_rate(case, p, act, D, t1, t2, vel, m, target_s: float, step0: int):
    """voxel-steps/s of the FP64 oracle over about target_s seconds of steps."""
    import oracle

    oracle.run(p, act, D, t1, t2, vel, 1, m, grad_sched=case.grad_sched[step0 : step0 + 1])  # warm
    t0 = time.perf_counter()
    oracle.run(p, act, D, t1, t2, vel, 1, m, grad_sched=case.grad_sched[step0 + 1 : step0 + 2])
    one = time.perf_counter() - t0
    n = max(1, min(int(target_s / max(one, 1e-6)), len(case.grad_sched) - step0 - 2))
    t0 = time.perf_counter()
    oracle.run(p, act, D, t1, t2, vel, n, m, grad_sched=case.grad_sched[step0 + 2 : step0 + 2 + n])
    dt = time.perf_counter() - t0
    return case.nvox * n / dt, n, dt


def cpu_baseline(case, target_s: float, scheme: int = 0):
    """FP64 oracle steps of the same grid on the host cores (bounded sample):
    all threads (value, cores) and one thread (single_core_value)."""
    import numpy as np

    import oracle

    p, act, D, t1, t2, vel, m = oracle.setup_case(case, m0=[case.mask.astype(np.float64),
                                                           np.zeros(case.shape), np.zeros(case.shape)],
                                                  scheme=scheme)
    threads = oracle.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    many, n_many, t_many = _oracle_rate(case, p, act, D, t1, t2, vel, m, target_s, 0)
    oracle.set_threads(1)
    one, n_one, t_one = _oracle_rate(case, p, act, D, t1, t2, vel, m, target_s / 2, 0)
    oracle.set_threads(threads)
    grid = "x".join(str(v) for v in case.shape[::-1])
    return {
        "value": many,
        "unit": "voxel-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"FP64 oracle (OpenMP restatement of {'ExplicitMakUpgrade' if scheme else 'ExplicitSimple'}), "
                  f"{grid}: {n_many} steps in {t_many:.1f} s on {threads} threads; {n_one} steps in "
                  f"{t_one:.1f} s on 1 thread",
        "single_core_value": one,
        "cpu_model": _cpu_model(),
    }


def copy_ceiling(mib: int = 1024, reps: int = 10) -> float:
    """GB/s of the library's own non-temporal dwordx4 copy kernel (bytes read
    plus bytes written over its event-timed duration)."""
    import ctypes

    from initializator_amd import _lib

    lib = _lib.load()
    ms = ctypes.c_double()
    nbytes = mib * (1 << 20)
    rc = lib.nmr_hip_copy_bandwidth(nbytes, reps, ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"nmr_hip_copy_bandwidth: {rc}")
    return 2 * nbytes / (ms.value * 1e-3) / 1e9


def mfma_ceiling(iters: int = 16384, reps: int = 5) -> float:
    """TFLOP/s of the library's matrix-core rate kernel: dense BF16
    v_mfma_f32_16x16x32_bf16 chains on register operands, four waves per
    SIMD, no memory traffic -- the ceiling the convolutions are priced at.
    16384 rounds (about 3.6 ms per launch) let the clocks settle: shorter
    launches read lower (256 rounds: about 2.1 PFLOP/s)."""
    import ctypes

    from initializator_amd import _lib

    lib = _lib.load()
    tf = ctypes.c_double()
    rc = lib.nmr_hip_mfma_rate(iters, reps, ctypes.byref(tf))
    if rc != 0:
        raise RuntimeError(f"nmr_hip_mfma_rate: {rc}")
    return tf.value


# MFMA work of the corrector's three layers per voxel: 35 instructions of
# 16 384 FLOP per 16-voxel row (7 for 6(->8)->16, 14 for 16->16, 14 for
# 16->2), padding included
CONV_MFMA_FLOP_PER_VOXEL = 35 * 16384 / 16


def copy_probe(mib: int) -> int:
    """The library's copy kernel on mib MiB: its GB/s is the streaming ceiling
    on this GPU, and its FETCH_SIZE / WRITE_SIZE counters (rocprofv3 --pmc)
    give the bytes-per-KiB scale of the sweep's counters."""
    gbs = copy_ceiling(mib, 10)
    print(json.dumps({"copy_mib": mib, "GBps": gbs, "ms": 2 * mib * (1 << 20) / (gbs * 1e9) * 1e3}), flush=True)
    return 0


def instance_errors(specs, k, nsteps, dx, dt, D):
    """Signal and field error against the FP64 oracle of the k smallest
    instances, advanced nsteps from their initial state in a batch of their
    own (a batched instance's result does not depend on its batch mates)."""
    import oracle
    from initializator_amd import generators
    from initializator_amd.solver import BatchSolver

    order = sorted(range(len(specs)), key=lambda i: int(np.prod(specs[i]["shape"])))[:k]
    sub = [specs[i] for i in order]
    b = BatchSolver(sub, 2, dx, dt, D, 1.0 / 2.0, r1_bulk=1.0 / 2.5e10, velocity=True, precession=True)
    for j, sp in enumerate(sub):
        b.upload(j, sp["mask"], [sp["mask"].astype(np.float64), np.zeros(sp["shape"])], sp["vel"])
    sig = b.run(nsteps)
    rows = []
    for j, sp in enumerate(sub):
        sched = generators.pgse_schedule(nsteps, dt, sp["grad_amp"], sp["delta"], sp["big_delta"])
        case = generators.Case(sp["shape"], sp["mask"], *sp["vel"], dx, dt, D, 2.0, 2.5e10, sp["rho1"], nsteps,
                               sched, 2)
        m0 = [sp["mask"].astype(np.float64), np.zeros(sp["shape"]), np.zeros(sp["shape"])]
        sig_ref, f_ref = oracle.run_case(case, m0=m0)
        f = b.download(j)
        vox = max(float(np.abs(f[q] - f_ref[q]).max()) for q in range(2))
        scale = max(float(np.abs(f_ref[q]).max()) for q in range(2))
        rows.append((float(np.linalg.norm(sig[j] - sig_ref[:, :2]) / np.linalg.norm(sig_ref[:, :2])), vox,
                     vox / max(scale, 1e-300)))
    b.close()
    e = np.array(rows)
    return {"instances": [list(specs[i]["shape"]) for i in order], "steps": nsteps,
            "signal_rel_l2_max": float(e[:, 0].max()), "max_voxel_max": float(e[:, 1].max()),
            # the linearised precession grows |M| under the strongest gradients
            # (b = 1e10 s/m^2 is 7.5 T/m): the field error relative to max |M|
            "max_voxel_rel_max": float(e[:, 2].max()),
            "signal_rel_l2": [float(x) for x in e[:, 0]]}


def batch_main(args) -> int:
    """Batched workload: many uneven instances advanced together."""
    from initializator_amd import generators
    from initializator_amd.solver import BatchSolver

    specs = generators.batch_instances(args.instances)
    print(f"[bench] generated {len(specs)} instances", file=sys.stderr, flush=True)
    dx, D = 1e-6, 2.3e-9
    dt = generators.stable_dt(dx, D)
    b = BatchSolver(specs, 2, dx, dt, D, 1.0 / 2.0, r1_bulk=1.0 / 2.5e10, velocity=True, precession=True)
    for i, sp in enumerate(specs):
        b.upload(i, sp["mask"], [sp["mask"].astype(np.float64), np.zeros(sp["shape"])], sp["vel"])
    if args.variant >= 0:
        b.set_tile(args.variant, args.lz, args.graph_steps, not args.full_grid)
        v, lz, per = args.variant, args.lz, not args.full_grid
    else:
        b.set_tile(-1, 0, args.graph_steps)
        v, lz, per = b.autotune(3)
        b.set_tile(v, lz, args.graph_steps, per)
    b.run(args.warmup, want_signal=False)
    t0 = time.perf_counter()
    b.run(args.steps, want_signal=True)
    wall = time.perf_counter() - t0
    gpu_ms, vox_steps = b.last_run_ms()
    b.close()
    nvox = sum(int(np.prod(sp["shape"])) for sp in specs)
    # the same instances solved one after another, one context each (the
    # unbatched path), on a shorter run: GPU time of their runs only
    loop = None
    if args.single_steps > 0:
        from initializator_amd.solver import HipSolver

        loop_ms = 0.0
        for sp in specs:
            h = HipSolver(sp["shape"], 2, dx, dt, D, 1.0 / 2.0, r1_bulk=1.0 / 2.5e10, rho1=sp["rho1"],
                          rho2=sp["rho2"], velocity=True, precession=True)
            h.upload_mask(sp["mask"])
            h.upload_velocity(*sp["vel"])
            h.upload_magnetization(sp["mask"].astype(np.float64), np.zeros(sp["shape"]))
            sched = generators.pgse_schedule(args.single_steps, dt, sp["grad_amp"], 2e-3, 20e-3)
            h.run(args.single_steps, sched)
            loop_ms += h.last_run_ms()[0]
            h.close()
        loop = {"value": nvox * args.single_steps / (loop_ms * 1e-3), "unit": "voxel-steps/s",
                "steps": args.single_steps, "gpu_ms": loop_ms}
    kernel_s = gpu_ms * 1e-3 / args.steps
    achieved = BYTES_PER_VOXEL_STEP * nvox / kernel_s / 1e9
    err = instance_errors(specs, args.error_instances, args.error_steps, dx, dt, D) if args.error_instances else None
    out = {
        "metric": "voxel-steps/s (batched uneven instances)",
        "value": nvox * args.steps / wall,
        "unit": "voxel-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded porosity, divergence-free velocity, PGSE per instance)",
        "config": {
            "workload": f"batched FP32 upwind Bloch-Torrey steps, {len(specs)} cubes of edge 32..128, "
                        "phi U(0.15,0.4), rho U(1,50) um/s, b log-U(1e7,1e10) s/m^2, complex (Mx,My)",
            "voxels": nvox,
            "variant": v,
            "lz": lz,
            "persistent_grid": per,
            "graph_steps": args.graph_steps,
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
        "single_instance_loop": loop,
        "per_instance_fp64_error": err,
    }
    print(json.dumps(out), flush=True)
    return 0


def learned_main(args) -> int:
    """Learned corrector at 256^3: train on the batched workload's instances,
    then the time to produce the signal curve with coarse steps of 8 dt
    (corrector, residual check, plain fallback) against the plain solve of
    the same problem, and the curves' relative L2 difference."""
    from initializator_amd import generators
    from initializator_amd.solver import BatchSolver, Corrector, HipSolver

    factor = 8
    t0 = time.perf_counter()
    specs = generators.batch_instances(args.instances)
    case = generators.case_flow(n=args.n, phi=args.phi, nsteps=args.steps)
    gen_s = time.perf_counter() - t0
    print(f"[bench] generated {len(specs)} training instances and the {args.n}^3 case in {gen_s:.1f} s",
          file=sys.stderr, flush=True)
    dx, D = 1e-6, 2.3e-9
    dt = generators.stable_dt(dx, D)
    # training on the batched generator's trajectories
    b = BatchSolver(specs, 2, dx, dt, D, 1.0 / 2.0, r1_bulk=1.0 / 2.5e10, velocity=True, precession=True,
                    phase_exact=True)
    for i, sp in enumerate(specs):
        b.upload(i, sp["mask"], [sp["mask"].astype(np.float64), np.zeros(sp["shape"])], sp["vel"])
    net = Corrector(args.seed)
    t0 = time.perf_counter()
    losses = net.train(b, factor=factor, rounds=args.train_rounds, stride=args.train_stride, lr=args.lr)
    train_s = time.perf_counter() - t0
    b.close()
    print(f"[bench] trained on {losses.size} samples in {train_s:.1f} s: loss {losses[:2].mean():.3e} -> "
          f"{losses[-2:].mean():.3e}", file=sys.stderr, flush=True)

    ncoarse = args.steps // factor
    nsteps = ncoarse * factor
    sched = case.grad_sched[:nsteps]

    def solver():
        s = HipSolver.from_case(case, phase_exact=True)
        s.upload_magnetization(case.mask.astype(np.float64), np.zeros(case.shape))
        return s

    # warm both paths once on a short run, then time each on a fresh state
    s = solver()
    s.run(2 * factor, sched[: 2 * factor])
    net.run_learned(s, 2, factor, sched[: 2 * factor], tol=args.tol)
    s.close()
    s = solver()
    t0 = time.perf_counter()
    plain = s.run(nsteps, sched)
    plain_s = time.perf_counter() - t0
    s.close()
    s = solver()
    t0 = time.perf_counter()
    sig, why = net.run_learned(s, ncoarse, factor, sched, tol=args.tol, reasons=True)
    nfall = why[0]
    learned_s = time.perf_counter() - t0
    s.close()
    ref = plain[::factor]
    err = float(np.linalg.norm(sig - ref) / np.linalg.norm(ref))
    net.close()
    # the floor the three layers' MFMA work puts under one coarse step, at
    # the measured matrix-core rate, against the 3x budget per coarse step
    mfma_tf = mfma_ceiling()
    conv_tflop = CONV_MFMA_FLOP_PER_VOXEL * case.nvox / 1e12
    budget_us = plain_s / ncoarse / 3 * 1e6
    out = {
        "metric": "learned vs plain time-to-signal @1e-3",
        "value": plain_s / learned_s,
        "unit": "x (plain wall-clock / learned wall-clock)",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": 2 * factor,
        "ms_per_step": learned_s * 1e3 / nsteps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 (net) / f32 (solver)",
        "data": "synthetic (training: batched generator instances; evaluation: seeded 256^3 porosity, "
                "divergence-free velocity, PGSE)",
        "config": {
            "workload": f"learned corrector, coarse steps of {factor} dt, {args.n}^3, phi={args.phi}, flow + PGSE "
                        "G=0.3 T/m, complex (Mx,My), exact phase",
            "plain_s": plain_s,
            "learned_s": learned_s,
            "signal_rel_l2": err,
            "within_1e-3": err <= 1e-3,
            "fallbacks": nfall,
            "fallbacks_nonfinite_energy_size": list(why[1:]),
            "coarse_steps": ncoarse,
            "tol": args.tol,
            "train_samples": int(losses.size),
            "train_s": train_s,
            "train_loss_first_last": [float(losses[:2].mean()), float(losses[-2:].mean())],
        },
        "mfma_floor": {
            "ceiling_tflops": mfma_tf,
            "peak_tflops": 2500.0,
            "conv_tflop_per_coarse_step": conv_tflop,
            "conv_floor_us": conv_tflop / mfma_tf * 1e6,
            "budget_us_per_coarse_step_at_3x": budget_us,
        },
    }
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=256, help="grid edge")
    ap.add_argument("--phi", type=float, default=0.20)
    ap.add_argument("--rho", type=float, default=10e-6, help="plain workload: surface relaxivity (m/s)")
    ap.add_argument("--variant", type=int, default=-1)
    ap.add_argument("--lz", type=int, default=0)
    ap.add_argument("--scheme", choices=("upwind", "maccormack"), default="upwind",
                    help="plain workload: ExplicitSimple's upwind step (default) or the reference GPU loop's "
                         "MacCormack step")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="CPU baseline sample: seconds of oracle steps on all threads (half that on one thread)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("plain", "batch", "learned"), default="plain",
                    help="plain: one 256^3 grid (default); batch: many uneven instances per launch; "
                         "learned: corrector training and learned-vs-plain time to signal")
    ap.add_argument("--train-rounds", type=int, default=40, help="learned workload: training rounds")
    ap.add_argument("--train-stride", type=int, default=120, help="learned workload: plain steps between rounds")
    ap.add_argument("--lr", type=float, default=1e-3, help="learned workload: Adam step size")
    ap.add_argument("--tol", type=float, default=1e-3, help="learned workload: residual tolerance")
    ap.add_argument("--seed", type=int, default=0, help="learned workload: weight seed")
    ap.add_argument("--instances", type=int, default=256, help="batch workload: number of instances")
    ap.add_argument("--graph-steps", type=int, default=64, help="batch workload: steps per captured graph")
    ap.add_argument("--full-grid", action="store_true",
                    help="batch workload with --variant: one workgroup per tile instead of a resident grid")
    ap.add_argument("--error-instances", type=int, default=8,
                    help="batch workload: smallest instances checked against the FP64 oracle (0: skip)")
    ap.add_argument("--error-steps", type=int, default=1000, help="batch workload: steps of that check")
    ap.add_argument("--shape", type=int, nargs=3, metavar=("NZ", "NY", "NX"), default=None,
                    help="plain workload: grid (default n^3)")
    ap.add_argument("--single-steps", type=int, default=200,
                    help="batch workload: steps of the one-instance-at-a-time comparison (0: skip)")
    ap.add_argument("--bc", choices=("neumann", "periodic", "dirichlet"), default="neumann",
                    help="plain workload: ghost rule on every axis (other than neumann: the general sweep)")
    ap.add_argument("--field", action="store_true",
                    help="plain workload: a seeded per-voxel field (+-20 uT) along z in place of the gradient "
                         "field's z term (the general sweep)")
    ap.add_argument("--coef", action="store_true",
                    help="plain workload: seeded per-voxel D and 1/T2 (the general sweep's coefficient variant)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="time a plain device-to-device copy of MIB MiB instead (HBM ceiling / counter scale)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU; the solve keeps the whole domain on one MI355X "
              "(multi-GPU z-slab sharding is a design note, not built)", file=sys.stderr)
        return 2
    if args.workload == "batch":
        return batch_main(args)
    if args.workload == "learned":
        return learned_main(args)

    from initializator_amd import generators
    from initializator_amd.solver import HipSolver

    from initializator_amd import _lib

    mak = args.scheme == "maccormack"
    bpv = BYTES_PER_VOXEL_STEP_MAK if mak else BYTES_PER_VOXEL_STEP
    # per-voxel inputs each pass reads: the field 4 B, D and 1/T2 8 B
    extra = (4.0 if args.field else 0.0) + (8.0 if args.coef else 0.0)
    bpv += extra * (2 if mak else 1)
    bc = {"neumann": _lib.NMR_BC_NEUMANN, "periodic": _lib.NMR_BC_PERIODIC,
          "dirichlet": _lib.NMR_BC_DIRICHLET}[args.bc]
    total = args.warmup + args.steps
    shape = tuple(args.shape) if args.shape else (args.n,) * 3
    case = generators.case_flow(shape=shape, phi=args.phi, nsteps=total, rho=args.rho,
                                progress=lambda msg: print(f"[bench] {msg}", file=sys.stderr, flush=True))
    # the streaming ceiling first: it also brings the GPU's clocks up before
    # the warm-up steps
    ceiling = copy_ceiling()
    s = HipSolver.from_case(case, scheme=_lib.NMR_SCHEME_MACCORMACK if mak else _lib.NMR_SCHEME_UPWIND,
                            bc=(bc, bc, bc))
    s.set_tile(args.variant, args.lz)
    rng = np.random.default_rng(0x4E4D52 + 7000)
    if args.field:
        s.upload_field(rng.uniform(-2e-5, 2e-5, case.shape))
    if args.coef:
        s.upload_coefficients(np.where(case.mask != 0, case.diffusivity, 0.0), rng.uniform(0.5, 5.0, case.shape))
    # the upload converts the state to the sweep's brick layout and sizes its
    # buffers; the state then stays on the device, in bricks, from run to run
    s.upload_magnetization(case.mask.astype("float64"), np.zeros(case.shape))
    # the warm-up takes the same path as the timed run (signal rows copied
    # back), so that no first call of it falls inside the timed region
    s.run(args.warmup, case.grad_sched[: args.warmup], want_signal=True)
    sched = case.grad_sched[args.warmup :]
    t0 = time.perf_counter()
    s.run(args.steps, sched, want_signal=True)
    wall = time.perf_counter() - t0
    # HIP events on the solver's stream around the timed run's kernels (the
    # sweep launches and one reduction of their partial sums)
    gpu_ms, nsw = s.last_run_ms()
    s.close()

    nvox = case.nvox
    value = nvox * args.steps / wall
    kernel_s = gpu_ms * 1e-3 / args.steps
    achieved = bpv * nvox / kernel_s / 1e9
    plain_physics = args.bc == "neumann" and not args.field and not args.coef
    default = (shape, args.phi, args.variant, args.lz, args.rho) == ((256, 256, 256), 0.20, -1, 0, 10e-6) \
        and plain_physics
    traffic = (MAK_TRAFFIC if mak else DEFAULT_TRAFFIC) if default else None
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
            "workload": f"plain FP32 {'MacCormack' if mak else 'upwind'} Bloch-Torrey step, "
                        f"{'x'.join(str(v) for v in shape[::-1])}, phi={args.phi}, rho={args.rho:g} m/s, "
                        "flow + PGSE G=0.3 T/m, complex (Mx,My)"
                        + ("" if args.bc == "neumann" else f", {args.bc} ghosts")
                        + (", per-voxel field along z" if args.field else "")
                        + (", per-voxel D and 1/T2" if args.coef else ""),
            "grid": list(shape[::-1]),
            "bytes_per_voxel_step": bpv,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            # traffic is not measured by this run: rocprofv3 --pmc FETCH_SIZE /
            # WRITE_SIZE of the same command, scaled by a plain copy's counters
            "traffic_source": (f"profiles/{'gsweep_maccormack' if mak else 'sweep_kernel'}_256cube_flow_pgse.md"
                               if traffic else None),
            # measured DRAM bytes over the same kernel time, against 8 TB/s
            "dram_frac": traffic / kernel_s / (HBM_PEAK_GBS * 1e9) if traffic else None,
            "kernel_us": kernel_s * 1e6,
            "sweeps_per_step": nsw / args.steps,
            # the library's own non-temporal 16-byte copy kernel on this GPU
            "copy_ceiling_GBps": ceiling,
            "frac_of_copy_ceiling": achieved / ceiling,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu and plain_physics:
        # the same grid and physics, with a schedule long enough for the sample
        cb = dataclasses.replace(case, nsteps=404, grad_sched=generators.pgse_schedule(404, case.dt, 0.3, 2e-3,
                                                                                       20e-3))
        out["cpu_baseline"] = cpu_baseline(cb, args.cpu_seconds, scheme=1 if mak else 0)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
