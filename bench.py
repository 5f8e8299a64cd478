"""Benchmark of the gfx950 LEQ_BICGS solver on one MI355X.

Workload (default): the 256^3 void-fraction pressure-correction system
(17,173,512 cells with ghosts; faces 10^U[-1,1] times the face-averaged gas
fraction of 2000 seeded spheres at 0.4 in a 1.0 background, B = U[-1e-3,1e-3],
seed 2000), DIAG preconditioner, zero initial guess, relative residual 1e-10,
solved by the mixed-precision solver (FP32 BiCGSTAB recurrence, FP64 true
residual, solution and scalars).  One step = one full solve from fresh
copies of A_M/B_M already resident in HBM (the solve scales them in place).

Printed: ONE JSON line with solves/s, p50/p99 solve latency, the plain FP64
solver on the same system, the HBM roofline of the dominant kernel (HIP
event times measured in this run) and a CPU baseline: the reference's own
LEQ_BICGS (compiled from the MFIX sources with flang -O2 -fopenmp) timed on
this host on a bounded sample of the same solve.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def _percentile(v, q):
    return float(np.percentile(np.asarray(v), q)) if len(v) else None


def cpu_baseline(a, b, n, its_per_solve, sample_its, threads):
    """Reference LEQ_BICGS (oracle/_ref) on `sample_its` iterations of the same
    system; solves/s extrapolated with the GPU FP64 iteration count."""
    import oracle
    if not oracle.REF_BIN.exists():
        return None
    g = oracle.grid(n, n, n)
    env_threads = str(threads)
    os.environ["OMP_NUM_THREADS"] = env_threads
    out = oracle.run_reference("bicgs", g, a, b, np.zeros(b.size), pc="DIAG", cmethod="RSRS", tol=1e-10,
                               itmax=sample_its, seed=2000)
    secs = float(out["dot"][1])
    its = out["iter_tot"] if out["ier"] == 0 else sample_its
    per_it = secs / max(1, sample_its)
    return {
        "value": 1.0 / (per_it * its_per_solve),
        "unit": "solves/s",
        "cores": threads,
        "kind": "reference",
        "sample": f"{sample_its} LEQ_BICGS iterations of the {n}^3 system ({secs:.1f} s, "
                  f"{per_it * 1e3:.0f} ms/iteration), extrapolated to the {its_per_solve} iterations "
                  f"of the FP64 solve",
        "ms_per_iteration": per_it * 1e3,
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256, help="interior cells per axis")
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--cpu-iters", type=int, default=6, help="reference iterations timed for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this solver has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        return 2

    from mfix_amd import LIB_PATH, leq, systems, _lib
    import ctypes as C
    if not LIB_PATH.exists():
        from mfix_amd import build
        build.build()
    leq.init(0)
    n = args.n
    t0 = time.time()
    a, b = systems.gen_system(systems.VOID_PRESSURE, n, n, n, 2000)
    gen_s = time.time() - t0
    n3 = b.size
    leq.grid_init(n, n, n)
    blk = C.c_int()
    lpb = C.c_int()
    for mixed in (1, 0):
        _lib.check(_lib.load().mfix_leq_autotune(n, n, n, mixed, C.byref(blk), C.byref(lpb)), "autotune")
    a0 = leq.DeviceArray.from_host(a)
    b0 = leq.DeviceArray.from_host(b)
    da, db = leq.DeviceArray(a.size), leq.DeviceArray(n3)
    dx = leq.DeviceArray(n3)
    zeros = np.zeros(n3)

    def one(mixed):
        da.copy_from(a0)
        db.copy_from(b0)
        dx.upload(zeros)
        t = time.perf_counter()
        out = leq.bicgs_dev(dx, da, db, pc="DIAG", cmethod="RSRS", tol=1e-10, itmax=20000, mixed=mixed)
        dt = time.perf_counter() - t
        if out.ier != 0:
            raise RuntimeError(f"solve failed: {out}")
        return dt, out

    for _ in range(args.warmup):
        one(True)
    _lib.check(_lib.load().mfix_leq_sync(), "sync")
    leq.prof_enable(0, True)
    times, stats = [], []
    t_begin = time.perf_counter()
    for _ in range(args.steps):
        dt, out = one(True)
        times.append(dt)
        stats.append(out.stats)
    _lib.check(_lib.load().mfix_leq_sync(), "sync")
    total = time.perf_counter() - t_begin
    prof = leq.prof_read(0)
    leq.prof_enable(0, False)

    fp64_times, fp64_stats = [], []
    for _ in range(args.fp64_steps):
        dt, out = one(False)
        fp64_times.append(dt)
        fp64_stats.append(out.stats)

    its = float(np.mean([s["iterations"] for s in stats]))
    fp64_its = float(np.mean([s["iterations"] for s in fp64_stats])) if fp64_stats else None
    value = args.steps / total
    fp64_value = len(fp64_times) / sum(fp64_times) if fp64_times else None

    # dominant kernel: the FP32 stencil kernel (V = A*Phat and T = A*Shat with
    # their dot products), 7 A32 planes + operand + dot partner + result = 40 B/cell
    kern_avg_s = prof["total_ms"] * 1e-3 / max(1, prof["launches"])
    kern_bytes = prof["alg_bytes"] / max(1, prof["launches"])
    achieved = kern_bytes / kern_avg_s if kern_avg_s > 0 else 0.0
    traffic = None
    tf = ROOT / "profiles" / "stencil_fp32_256cubed_traffic.json"
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get("bytes_per_launch")
        except Exception:
            traffic = None
    # whole-solve algorithmic rate: iterations x fused-schedule bytes per cell
    # (DIAG: 15 sA + 15 sV + 2 sX with sA = sV = 4, sX = 8 -> 136 B/cell)
    solve_bytes = its * n3 * 136.0
    solve_rate = solve_bytes / (total / args.steps)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(a, b, n, int(round(fp64_its or its)), args.cpu_iters,
                               int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        except Exception as e:  # the baseline is reported, never required
            print(f"cpu baseline skipped: {e}", file=sys.stderr)

    line = {
        "metric": "solves/s (FP32-refined, 1e-10 relative residual)",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"void-fraction pressure-correction {n}^3 (N3={n3}), DIAG, x0=0, TOL=1e-10, "
                        "mixed FP32/FP64 LEQ_BICGS",
            "n3": n3,
            "iterations": its,
            "reliable_updates": float(np.mean([s["refinements"] for s in stats])),
            "latency_ms_p50": 1e3 * _percentile(times, 50),
            "latency_ms_p99": 1e3 * _percentile(times, 99),
            "fp64_solves_per_s": fp64_value,
            "fp64_iterations": fp64_its,
            "speedup_vs_fp64": (value / fp64_value) if fp64_value else None,
            "solve_alg_GBps": solve_rate / 1e9,
            "solve_roofline_frac": solve_rate / HBM_PEAK,
            "true_rel_residual": float(np.max([s["true_rnorm"] / s["rnorm0"] for s in stats])),
            "generation_s": gen_s,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_matvec_dots<float> (FP32 stencil + dots)",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic,
            "launches": prof["launches"],
            "avg_launch_us": kern_avg_s * 1e6,
            "alg_bytes_per_launch": kern_bytes,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
