"""Benchmark of the gfx950 LEQ_BICGS solver on one MI355X.

Workload (default): the 256^3 void-fraction pressure-correction system
(17,173,512 cells with ghosts; faces 10^U[-1,1] times the face-averaged gas
fraction of 2000 seeded spheres at 0.4 in a 1.0 background, B = U[-1e-3,1e-3],
seed 2000), DIAG preconditioner, zero initial guess, relative residual 1e-10,
solved by the mixed-precision solver (FP32 BiCGSTAB recurrence, FP64 true
residual, solution and scalars).  One step = one full solve from fresh
copies of A_M/B_M already resident in HBM (the solve scales them in place).

Printed: ONE JSON line with solves/s, p50/p99 solve latency, the plain FP64
solver on the same system and the mixed result's error against it, the HBM
roofline of the dominant kernel (kernel start and end times the dispatch
records in HIP events, measured in this run, over the launches that did
work) and a CPU baseline: the oracle's C/OpenMP
restatement of the reference's LEQ_BICGS (its loop structure and serial dot
products) timed on this host's cores on a bounded sample of the same solve.
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
    """The oracle's LEQ_BICGS (oracle/leq_oracle.c: the reference's algorithm
    and OpenMP loop structure, serial IS_SERIAL dot products) on this host's
    cores, on a bounded sample of the same FP64 DIAG solve: the set-up
    (unit scaling, R, R~) and `sample_its` iterations, timed inside the
    call; solves/s = 1 / (set-up + iterations of the full solve x median
    iteration time), with the GPU FP64 solve's iteration count."""
    import oracle
    prev = oracle.set_threads(threads)
    try:
        g = oracle.grid(n, n, n)
        noise = np.random.default_rng(2017).random(b.size)
        out = oracle.leq_bicgs(g, "DIAG", "RSRS", 1e-10, sample_its, np.zeros(b.size), a, b, noise, timing=True)
    finally:
        oracle.set_threads(prev)
    res, tt = out[3], out[4]
    ends = tt[np.isfinite(tt)]
    per_it = np.diff(ends)[2:] if ends.size > 4 else np.diff(ends)  # skip the first iterations (page faults)
    per_it_s = float(np.median(per_it))
    setup_s = float(res["seconds_setup"])
    return {
        "value": 1.0 / (setup_s + per_it_s * its_per_solve),
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle LEQ_BICGS (C/OpenMP restatement of model/leq_bicgs.f, {threads} threads) on the same "
                  f"{n}^3 system: set-up {setup_s:.2f} s + median of {per_it.size} iteration times "
                  f"({per_it_s * 1e3:.1f} ms) x the {its_per_solve} iterations of the FP64 solve "
                  f"({res['seconds_total']:.1f} s sampled)",
        "ms_per_iteration": per_it_s * 1e3,
        "setup_s": setup_s,
    }


def _bytes_per_cell(pc, mixed):
    """Fused-schedule algorithmic bytes per cell per iteration (sA, sV, sX =
    coefficient, Krylov-vector and solution element sizes)."""
    sa, sv, sx = (4, 4, 8) if mixed else (8, 8, 8)
    return {"DIAG": 15 * sa + 15 * sv + 2 * sx, "NONE": 14 * sa + 15 * sv + 2 * sx,
            "LINE": 42 * sa + 30 * sv + 2 * sx}[pc]


def bench_batch(args):
    """64 independent systems, sides log-uniform in [16,128], half symmetric
    pressure-correction and half upwind momentum, DIAG or LINE per system,
    solved together by mfix_leq_batch_solve (one launch per phase for all
    systems, iterations replayed from captured HIP graphs).  A step is one
    batched solve of all 64 from fresh resident copies of A_M/B_M (the copy
    refresh is outside the timed call)."""
    from mfix_amd import leq, systems, _lib
    cfg = systems.batch_config(64, seed=3000)
    held = []
    for c in cfg:
        side = c["side"]
        a, b = systems.gen_system(c["kind"], side, side, side, c["seed"])
        held.append(dict(c=c, a0=leq.DeviceArray.from_host(a), b0=leq.DeviceArray.from_host(b),
                         da=leq.DeviceArray(a.size), db=leq.DeviceArray(b.size), dx=leq.DeviceArray(b.size),
                         zeros=np.zeros(b.size), n3=b.size))

    def run(mixed):
        items = []
        for h in held:
            h["da"].copy_from(h["a0"])
            h["db"].copy_from(h["b0"])
            h["dx"].upload(h["zeros"])
            c = h["c"]
            items.append(leq.BatchSystem(c["side"], c["side"], c["side"], h["dx"], h["da"], h["db"], pc=c["pc"],
                                         cmethod="RSRS", tol=1e-10, itmax=5000, mixed=mixed, seed=c["seed"]))
        _lib.check(_lib.load().mfix_leq_sync(), "sync")
        t = time.perf_counter()
        res = leq.batch_solve(items)
        dt = time.perf_counter() - t
        bad = [q for q, r in enumerate(res) if r.ier != 0]
        if bad:
            raise RuntimeError(f"batched systems {bad} did not converge: {[res[q].stats for q in bad]}")
        return dt, res

    out = {}
    for mixed in (True, False):
        for _ in range(args.warmup):
            run(mixed)
        times, alg = [], []
        for _ in range(args.steps if mixed else max(1, args.fp64_steps)):
            dt, res = run(mixed)
            times.append(dt)
            alg.append(sum(r.stats["iterations"] * h["n3"] * _bytes_per_cell(h["c"]["pc"], mixed)
                           for r, h in zip(res, held)))
        total = sum(times)
        out[mixed] = dict(value=len(held) * len(times) / total, ms=1e3 * total / len(times),
                          alg_rate=sum(alg) / total, iters=[r.stats["iterations"] for r in res],
                          true_rel=max((r.stats["true_rnorm"] / r.stats["rnorm0"]) for r in res) if mixed else None,
                          lat=times)
    m, f = out[True], out[False]
    sides = [h["c"]["side"] for h in held]
    line = {
        "metric": "solves/s (batched, FP32-refined, 1e-10 relative residual)",
        "value": m["value"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": m["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {
            "workload": "64 systems, sides log-uniform 16..128 (seed 3000), 50% symmetric pressure / 50% upwind "
                        "momentum, DIAG or LINE(RSRS) per system, TOL=1e-10, one batched solve per step",
            "sides_min_median_max": [min(sides), float(np.median(sides)), max(sides)],
            "total_cells": int(sum(h["n3"] for h in held)),
            "iterations_sum": int(sum(m["iters"])),
            "fp64_solves_per_s": f["value"], "fp64_iterations_sum": int(sum(f["iters"])),
            "speedup_vs_fp64": m["value"] / f["value"],
            "solve_alg_GBps": m["alg_rate"] / 1e9, "solve_roofline_frac": m["alg_rate"] / HBM_PEAK,
            "fp64_alg_GBps": f["alg_rate"] / 1e9, "fp64_roofline_frac": f["alg_rate"] / HBM_PEAK,
            "max_true_rel_residual": m["true_rel"],
        },
        "roofline": {"bound": "hbm", "kernel": "whole batched solve (fused-schedule bytes)",
                     "achieved": m["alg_rate"] / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": m["alg_rate"] / HBM_PEAK, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(line))
    return 0


def bench_line(args):
    """Plain FP64 LEQ_BICGS with the LINE preconditioner (RSRS) on the 128^3
    first-order-upwind momentum system (seed 1000), the reference's own
    arithmetic; also the mixed-precision solve of the same system."""
    from mfix_amd import leq, systems, _lib
    n = 128
    a, b = systems.gen_system(systems.MOMENTUM, n, n, n, 1000)
    leq.grid_init(n, n, n)
    a0, b0 = leq.DeviceArray.from_host(a), leq.DeviceArray.from_host(b)
    da, db, dx = leq.DeviceArray(a.size), leq.DeviceArray(b.size), leq.DeviceArray(b.size)
    zeros = np.zeros(b.size)

    def one(mixed):
        da.copy_from(a0)
        db.copy_from(b0)
        dx.upload(zeros)
        _lib.check(_lib.load().mfix_leq_sync(), "sync")
        t = time.perf_counter()
        out = leq.bicgs_dev(dx, da, db, pc="LINE", cmethod="RSRS", tol=1e-10, itmax=500, mixed=mixed)
        dt = time.perf_counter() - t
        if out.ier != 0:
            raise RuntimeError(f"solve failed: {out}")
        return dt, out

    res = {}
    for mixed in (False, True):
        for _ in range(args.warmup):
            one(mixed)
        leq.prof_enable(1, True)
        times, its = [], []
        for _ in range(args.steps):
            dt, out = one(mixed)
            times.append(dt)
            its.append(out.stats["iterations"])
        pr = leq.prof_read(1)
        leq.prof_enable(1, False)
        res[mixed] = dict(value=len(times) / sum(times), its=float(np.mean(its)), ms=1e3 * sum(times) / len(times),
                          line_us=1e3 * pr["total_ms"] / max(1, pr["launches"]),
                          line_GBps=pr["alg_bytes"] / max(1e-12, pr["total_ms"] * 1e-3) / 1e9)
    f, m = res[False], res[True]
    n3 = b.size
    alg = f["its"] * n3 * _bytes_per_cell("LINE", False) / (f["ms"] * 1e-3)
    line = {
        "metric": "solves/s (FP64 LEQ_BICGS, LINE/RSRS, 1e-10 relative residual)",
        "value": f["value"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": f["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"upwind momentum {n}^3 (N3={n3}), LINE RSRS, x0=0, TOL=1e-10, ITMAX=500",
                   "iterations": f["its"], "line_phase_us": f["line_us"], "line_phase_alg_GBps": f["line_GBps"],
                   "mixed_solves_per_s": m["value"], "mixed_iterations": m["its"],
                   "solve_alg_GBps": alg / 1e9, "solve_roofline_frac": alg / HBM_PEAK},
        "roofline": {"bound": "hbm", "kernel": "k_line_color_sys<double> (one red-black colour of line solves)",
                     "achieved": f["line_GBps"], "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": f["line_GBps"] * 1e9 / HBM_PEAK, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(line))
    return 0


def bench_resid(args):
    """CALC_RESID_C / _S / _PP (model/calc_resid.f) on the 256^3 void-fraction
    pressure system (seed 2000): ghost cells walls (FLAG 100), interior fluid
    (FLAG 1), VAR uniform in [-1,1].  One call = the residual pass, the
    finish block and the copy of the result to the host (blocking)."""
    from mfix_amd import leq, systems, _lib
    n = args.n
    a, b = systems.gen_system(systems.VOID_PRESSURE, n, n, n, 2000)
    n3 = b.size
    leq.grid_init(n, n, n)
    f = np.full((n + 2, n + 2, n + 2), 100, dtype=np.int32)
    f[1:-1, 1:-1, 1:-1] = 1
    x = np.random.default_rng(5).uniform(-1, 1, n3)
    da, db, dx = leq.DeviceArray.from_host(a), leq.DeviceArray.from_host(b), leq.DeviceArray.from_host(x)
    dflag = leq.DeviceArray.from_host(f.reshape(-1))
    per_cell = {"C": 7 * 8 + 8 + 8 + 4, "S": 7 * 8 + 8 + 8 + 4, "PP": 8 + 4}
    res = {}
    for kind in ("C", "S", "PP"):
        def one():
            return leq.calc_resid_dev(kind, b=db, flag=dflag, var=dx, a=da, tol=0.0, norm=1.0, debug=True)
        for _ in range(max(1, args.warmup)):
            one()
        t = time.perf_counter()
        for _ in range(args.steps):
            out = one()
        dt = (time.perf_counter() - t) / args.steps
        res[kind] = dict(us=dt * 1e6, GBps=per_cell[kind] * n3 / dt / 1e9, resid=out["resid"],
                         ncells=out["ncells"])
    c = res["C"]
    line = {
        "metric": "CALC_RESID_C calls/s (residual norm of an assembled equation)",
        "value": 1e6 / c["us"], "unit": "calls/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": c["us"] / 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"void-fraction pressure {n}^3 (N3={n3}), ghost walls, fluid interior",
                   "per_kind": res},
        "roofline": {"bound": "hbm", "kernel": "k_resid (CALC_RESID_C) + finish + result copy",
                     "achieved": c["GBps"], "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": c["GBps"] * 1e9 / HBM_PEAK, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(line))
    return 0


def bench_pp(args):
    """The gas pressure-correction equation of one SIMPLE iteration assembled
    and solved on the device (mfix_solve_pp_g_dev: INIT_AB_M + CONV_PP_G +
    SOURCE_PP_G, the residual normalisation and LEQ_BICGS with MFIX's
    defaults for equation 1: LINE, RSRS, TOL 1e-4, 20 iterations, FP64) on
    seeded fields of an n^3 box with a pressure-outflow face.  Beside it,
    the host-array path a CPU-assembled A_M/B_M takes (mfix_leq_bicgs: A_M,
    B_M and PP_G over PCIe, the same solve, PP_G back; the CPU assembly
    itself is not timed)."""
    from mfix_amd import leq, systems, _lib
    n = args.n
    fh, flag_h = systems.pp_g_fields(n, n, n, 2024)
    n3 = flag_h.size
    leq.grid_init(n, n, n)
    fields = {k: leq.DeviceArray.from_host(v) for k, v in fh.items()}
    flag = leq.DeviceArray.from_host(flag_h)
    da, db, dbm, dpp = leq.DeviceArray(7 * n3), leq.DeviceArray(n3), leq.DeviceArray(n3), leq.DeviceArray(n3)
    L = _lib.load()
    for _ in range(3):
        leq.assemble_pp_g_dev(fields, flag, 100.0, 0, da, db, dbm)
    _lib.check(L.mfix_leq_sync(), "sync")
    reps = 20
    t = time.perf_counter()
    for _ in range(reps):
        leq.assemble_pp_g_dev(fields, flag, 100.0, 0, da, db, dbm)
    asm_s = (time.perf_counter() - t) / reps
    asm_bytes = (16 * 8 + 4 + 7 * 8 + 8 + 8) * n3
    kw = dict(itmax=20, sweep="RSRS", tol=1e-4, pc="LINE", mixed=False)
    for _ in range(args.warmup):
        leq.solve_pp_g_dev(fields, flag, 100.0, 0, 0.0, dpp, da, db, dbm, **kw)
    times = []
    for _ in range(args.steps):
        _lib.check(L.mfix_leq_sync(), "sync")
        t = time.perf_counter()
        out = leq.solve_pp_g_dev(fields, flag, 100.0, 0, 0.0, dpp, da, db, dbm, **kw)
        times.append(time.perf_counter() - t)
    dev_ms = 1e3 * sum(times) / len(times)
    # host path: the assembled system as a CPU assembly would hand it over
    leq.assemble_pp_g_dev(fields, flag, 100.0, 0, da, db, dbm)
    a_h, b_h = da.download(), db.download()
    _lib.check(L.mfix_leq_set_write_back(0), "write_back")
    host_times = []
    for _ in range(args.steps):
        a_c, b_c, x_c = a_h.copy(), b_h.copy(), np.zeros(n3)
        t = time.perf_counter()
        ier_h = leq.leq_bicgs("Pp_g", 1, x_c, a_c, b_c, "RSRS", 1e-4, "LINE", 20)
        host_times.append(time.perf_counter() - t)
    _lib.check(L.mfix_leq_set_write_back(1), "write_back")
    host_ms = 1e3 * sum(host_times) / len(host_times)
    line = {
        "metric": "SOLVE_PP_G/s (gas pressure correction assembled and solved on the device)",
        "value": 1e3 / dev_ms, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dev_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"gas pressure correction {n}^3 (N3={n3}), pressure-outflow face, "
                               "LINE RSRS, TOL 1e-4, ITMAX 20 (MFIX defaults, equation 1)",
                   "iterations": out["stats"]["iterations"], "ier": out["ier"], "resg": out["resg"],
                   "assembly_us": asm_s * 1e6, "assembly_alg_GBps": asm_bytes / asm_s / 1e9,
                   "host_path_ms": host_ms, "host_path_ier": ier_h,
                   "host_copy_bytes_avoided": 9 * 8 * n3},
        "roofline": {"bound": "hbm", "kernel": "k_assemble_pp_g (+ counter read-back)",
                     "achieved": asm_bytes / asm_s / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": asm_bytes / asm_s / HBM_PEAK, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(line))
    return 0


def bench_phi(args):
    """A gas-phase scalar transport equation of one SIMPLE iteration on the
    device, in SOLVE_SCALAR_EQ's order (model/solve_scalar_eq.f): INIT_AB_M +
    CONV_DIF_PHI (first-order upwinding), SOURCE_PHI, CALC_RESID_S,
    UNDER_RELAX_S (UR_FAC(9) = 0.8) and LEQ_BICGS with MFIX's defaults for
    equation 9 (LINE, RSRS, TOL 1e-4, 15 iterations, FP64), on seeded fields
    of an n^3 box (systems.phi_fields).  BC_PHI is not applied (it stays with
    the caller).  Each kernel is also timed alone; beside the chain, the
    host-array path a CPU-assembled A_M/B_M takes (A_M, B_M and PHI over PCIe,
    the same solve, PHI back)."""
    from mfix_amd import leq, systems, _lib
    n = args.n
    fh, flag_h = systems.phi_fields(n, n, n, 2025)
    n3 = flag_h.size
    leq.grid_init(n, n, n)
    fields = {k: leq.DeviceArray.from_host(v) for k, v in fh.items()}
    flag = leq.DeviceArray.from_host(flag_h)
    da, db = leq.DeviceArray(7 * n3), leq.DeviceArray(n3)
    L = _lib.load()
    src = (fields["s_p"], fields["s_c"], fields["ep"], fields["phi"])
    stages = {
        "conv_dif_phi": (lambda: leq.conv_dif_phi_dev(fields, flag, da, db, sync=False), 148),
        "source_phi": (lambda: leq.source_phi_dev(*src, flag, da, db, sync=False), 108),
        "under_relax_s": (lambda: leq.under_relax_s_dev(fields["phi"], flag, 1.0, da, db, sync=False), 44),
    }
    per = {}
    reps = 20
    for name, (fn, bpc) in stages.items():
        for _ in range(3):
            fn()
        _lib.check(L.mfix_leq_sync(), "sync")
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        _lib.check(L.mfix_leq_sync(), "sync")
        sec = (time.perf_counter() - t) / reps
        per[name] = {"us": sec * 1e6, "bytes_per_cell": bpc, "alg_GBps": bpc * n3 / sec / 1e9,
                     "frac": bpc * n3 / sec / HBM_PEAK}
    kw = dict(pc="LINE", cmethod="RSRS", tol=1e-4, itmax=15)
    phi0 = fh["phi"]

    def chain():
        leq.conv_dif_phi_dev(fields, flag, da, db, sync=False)
        leq.source_phi_dev(*src, flag, da, db, sync=False)
        r = leq.calc_resid_dev("S", var=fields["phi"], a=da, b=db, flag=flag, debug=False)
        leq.under_relax_s_dev(fields["phi"], flag, 0.8, da, db, sync=False)
        return r, leq.bicgs_dev(fields["phi"], da, db, **kw)
    for _ in range(args.warmup):
        fields["phi"].upload(phi0)
        chain()
    times = []
    for _ in range(args.steps):
        fields["phi"].upload(phi0)
        _lib.check(L.mfix_leq_sync(), "sync")
        t = time.perf_counter()
        r, out = chain()
        times.append(time.perf_counter() - t)
    dev_ms = 1e3 * sum(times) / len(times)

    # the same chain with CALC_RESID_S and UNDER_RELAX_S done by the solver's
    # set-up pass (mfix_leq_resid_relax_bicgs_dev)
    def fused_chain():
        leq.conv_dif_phi_dev(fields, flag, da, db, sync=False)
        leq.source_phi_dev(*src, flag, da, db, sync=False)
        return leq.resid_relax_bicgs_dev("S", fields["phi"], da, db, flag, 0.8, debug=False, **kw)
    fused_times = []
    for _ in range(args.steps + args.warmup):
        fields["phi"].upload(phi0)
        _lib.check(L.mfix_leq_sync(), "sync")
        t = time.perf_counter()
        fr, fout = fused_chain()
        fused_times.append(time.perf_counter() - t)
    fused_times = fused_times[args.warmup:]
    fused_ms = 1e3 * sum(fused_times) / len(fused_times)
    if fout.stats["iterations"] != out.stats["iterations"] or fr["ncells"] != r["ncells"]:
        raise RuntimeError(f"fused chain differs: {fout.stats} {fr} against {out.stats} {r}")
    # host path: the assembled, under-relaxed system as a CPU assembly would hand it over
    fields["phi"].upload(phi0)
    leq.conv_dif_phi_dev(fields, flag, da, db)
    leq.source_phi_dev(*src, flag, da, db)
    leq.under_relax_s_dev(fields["phi"], flag, 0.8, da, db)
    a_h, b_h = da.download(), db.download()
    _lib.check(L.mfix_leq_set_write_back(0), "write_back")
    host_times = []
    for _ in range(args.steps):
        a_c, b_c, x_c = a_h.copy(), b_h.copy(), phi0.copy()
        t = time.perf_counter()
        ier_h = leq.leq_bicgs("Scalar 1", 9, x_c, a_c, b_c, "RSRS", 1e-4, "LINE", 15)
        host_times.append(time.perf_counter() - t)
    _lib.check(L.mfix_leq_set_write_back(1), "write_back")
    host_ms = 1e3 * sum(host_times) / len(host_times)
    conv = per["conv_dif_phi"]
    line = {
        "metric": "scalar transport equations/s (assembled, under-relaxed and solved on the device)",
        "value": 1e3 / dev_ms, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dev_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"gas-phase scalar {n}^3 (N3={n3}), first-order upwinding, UR 0.8, "
                               "LINE RSRS, TOL 1e-4, ITMAX 15 (MFIX defaults, equation 9)",
                   "iterations": out.stats["iterations"], "ier": out.ier, "resid_num": r["num"],
                   "resid_den": r["den"], "kernels": per, "host_path_ms": host_ms, "host_path_ier": ier_h,
                   "fused_resid_relax_chain_ms": fused_ms,
                   "host_copy_bytes_avoided": 10 * 8 * n3},
        "roofline": {"bound": "hbm", "kernel": "k_conv_dif_phi", "achieved": conv["alg_GBps"],
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": conv["frac"], "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(line))
    return 0


def bench_sweep(args):
    """One SOLVE_VEL_STAR stage of a SIMPLE iteration (model/solve_vel_star.f):
    U, V and W of the gas and two solids phases, nine upwind momentum systems
    on one n^3 grid that do not depend on each other's solutions, with MFIX's
    momentum defaults (LINE, RSRS, TOL 1e-4, LEQ_IT 5, FP64).  The stage is
    timed as nine sequential device solves (SOLVE_LIN_EQ call by call) and as
    nine deferred calls solved by one sweep flush."""
    from mfix_amd import leq, systems, _lib
    n = args.n
    leq.grid_init(n, n, n)
    L = _lib.load()
    names = [(f"{c}_{p}", vno) for p in ("g", "s1", "s2") for c, vno in (("U", 3), ("V", 4), ("W", 5))]
    src = [systems.gen_system(systems.MOMENTUM, n, n, n, 6000 + q, nspheres=6) for q in range(len(names))]
    n3 = src[0][1].size
    dev = [(leq.DeviceArray(7 * n3), leq.DeviceArray(n3), leq.DeviceArray(n3)) for _ in names]
    kw = dict(pc="LINE", cmethod="RSRS", tol=1e-4, itmax=5)
    zeros = np.zeros(n3)

    def load():
        for (a, b), (da, db, dx) in zip(src, dev):
            da.upload(a)
            db.upload(b)
            dx.upload(zeros)
        _lib.check(L.mfix_leq_sync(), "sync")

    def sequential():
        return [leq.bicgs_dev(dx, da, db, **kw) for da, db, dx in dev]

    def batched():
        for (name, vno), (da, db, dx) in zip(names, dev):
            leq.solve_lin_eq_deferred(name, vno, dx, da, db, kw["itmax"], kw["cmethod"], kw["tol"], kw["pc"])
        return leq.sweep_flush()

    times = {}
    for label, fn in (("sequential", sequential), ("batched", batched)):
        for _ in range(args.warmup):
            load()
            fn()
        ts = []
        for _ in range(args.steps):
            load()
            t = time.perf_counter()
            out = fn()
            ts.append(time.perf_counter() - t)
        times[label] = (1e3 * sum(ts) / len(ts), out)
    seq_ms, seq_out = times["sequential"]
    bat_ms, bat_out = times["batched"]
    iters = [d.stats["iterations"] for d in bat_out]
    assert iters == [r.stats["iterations"] for r in seq_out]
    line = {
        "metric": "SOLVE_VEL_STAR stages/s (nine momentum equations of one SIMPLE iteration, batched)",
        "value": 1e3 / bat_ms, "unit": "stages/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": bat_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"U, V, W of the gas and two solids phases, upwind momentum {n}^3 (N3={n3}), "
                               "LINE RSRS, TOL 1e-4, ITMAX 5 (MFIX momentum defaults)",
                   "equations": len(names), "iterations": iters, "sequential_ms": seq_ms, "batched_ms": bat_ms,
                   "speedup": seq_ms / bat_ms},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line))
    return 0


def bench_serve(args):
    """Poisson request streams through the continuous-batching server: sides
    lognormal (median 64, sigma 0.5) clipped to [24,160], 60% upwind momentum
    (LINE) / 40% symmetric pressure (DIAG), TOL 1e-8 or 1e-10.  Latency is
    arrival -> solution in the caller's host VAR; throughput is completed
    requests over the span from the first arrival to the last completion."""
    from mfix_amd import leq, systems
    from mfix_amd.serve import Server
    rates = [float(r) for r in args.rates.split(",") if r]
    nreq = args.requests
    proto = systems.serving_requests(rates[0], nreq, seed=4000)
    arrays = []
    for r in proto:
        side = r["side"]
        a, b = systems.gen_system(r["kind"], side, side, side, r["seed"])
        arrays.append((a, b))
    results = []
    for mixed in (True, False):
        for rate in rates:
            reqs = systems.serving_requests(rate, nreq, seed=4000)
            with Server(64 << 30, args.slots) as srv:
                live = {}
                done_t = {}
                nxt = 0
                t0 = time.perf_counter()
                while len(done_t) < nreq:
                    now = time.perf_counter() - t0
                    while nxt < nreq and reqs[nxt]["t"] <= now:
                        r = reqs[nxt]
                        a, b = arrays[nxt]
                        pc = "LINE" if r["kind"] == systems.MOMENTUM else "DIAG"
                        rid = srv.submit(a, b, np.zeros(b.size), r["side"], r["side"], r["side"],
                                         pc=pc, tol=r["tol"], itmax=5000, mixed=mixed, seed=r["seed"])
                        live[rid] = nxt
                        nxt += 1
                    if srv.pending():
                        srv.step()
                    elif nxt < nreq:
                        time.sleep(max(0.0, min(0.001, reqs[nxt]["t"] - (time.perf_counter() - t0))))
                    for rid in list(live):
                        d = srv.poll(rid)
                        if d is not None:
                            if d.ier != 0:
                                raise RuntimeError(f"request {live[rid]} failed: {d.stats}")
                            done_t[live.pop(rid)] = time.perf_counter() - t0
                lat = [done_t[q] - reqs[q]["t"] for q in range(nreq)]
                span = max(done_t.values()) - reqs[0]["t"]
                results.append(dict(mixed=mixed, rate=rate, p50_ms=1e3 * _percentile(lat, 50),
                                    p99_ms=1e3 * _percentile(lat, 99), solves_per_s=nreq / span))
                print(f"serve mixed={mixed} rate={rate:g}/s p50={results[-1]['p50_ms']:.1f} ms "
                      f"p99={results[-1]['p99_ms']:.1f} ms {results[-1]['solves_per_s']:.1f} solves/s",
                      file=sys.stderr, flush=True)
    top = max((r for r in results if r["mixed"]), key=lambda r: r["rate"])
    line = {
        "metric": "p50/p99 latency and solves/s under Poisson load (continuous batching)",
        "value": top["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": nreq, "warmup": 0,
        "ms_per_step": 1e3 / top["solves_per_s"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"{nreq} requests per rate, sides lognormal(64, 0.5) clipped to [24,160], 60% "
                               f"upwind LINE / 40% pressure DIAG, TOL 1e-8|1e-10, HBM budget 64 GiB, {args.slots} slots",
                   "p50_ms": top["p50_ms"], "p99_ms": top["p99_ms"], "rate": top["rate"], "sweep": results},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line))
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256, help="interior cells per axis")
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--cpu-iters", type=int, default=40, help="iterations of the oracle timed for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--workload", choices=["mixed", "batch", "serve", "line", "resid", "pp", "phi", "sweep"], default="mixed",
                    help="mixed: the 256^3 solve (headline); batch: 64 systems 16^3..128^3 in one batched "
                         "solve; serve: Poisson request streams through the continuous-batching server")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run the plain copy kernel 3x on 512 MiB (HBM counter calibration)")
    ap.add_argument("--rates", default="50,200,1000,2000", help="serve: request rates (1/s)")
    ap.add_argument("--requests", type=int, default=120, help="serve: requests per rate")
    ap.add_argument("--slots", type=int, default=64, help="serve: server slots (systems iterated together)")
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
    if args.copy_calibration:
        src = leq.DeviceArray(64 << 20)
        dst = leq.DeviceArray(64 << 20)
        for _ in range(3):
            _lib.check(_lib.load().mfix_leq_copy_kernel_dev(dst.ptr, src.ptr, 64 << 20), "copy kernel")
        src.free()
        dst.free()
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "line":
        return bench_line(args)
    if args.workload == "resid":
        return bench_resid(args)
    if args.workload == "pp":
        return bench_pp(args)
    if args.workload == "phi":
        return bench_phi(args)
    if args.workload == "sweep":
        return bench_sweep(args)
    n = args.n
    t0 = time.time()
    a, b = systems.gen_system(systems.VOID_PRESSURE, n, n, n, 2000)
    gen_s = time.time() - t0
    n3 = b.size
    leq.grid_init(n, n, n)
    blk = C.c_int()
    lpb = C.c_int()
    shapes = {}
    for mixed in ((1, 0) if not args.no_tune else ()):
        _lib.check(_lib.load().mfix_leq_autotune(n, n, n, mixed, C.byref(blk), C.byref(lpb)), "autotune")
        shapes["mixed" if mixed else "fp64"] = f"{blk.value // 100}x{blk.value % 100}"
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

    x_mixed = dx.download()
    fp64_times, fp64_stats = [], []
    x_fp64 = None
    for _ in range(args.fp64_steps):
        dt, out = one(False)
        fp64_times.append(dt)
        fp64_stats.append(out.stats)
        x_fp64 = dx.download()
    # error of the mixed-precision result against the FP64 solve (max norm,
    # relative to the FP64 solution's max norm)
    err_vs_fp64 = (float(np.max(np.abs(x_mixed - x_fp64)) / np.max(np.abs(x_fp64)))
                   if x_fp64 is not None else None)

    its = float(np.mean([s["iterations"] for s in stats]))
    fp64_its = float(np.mean([s["iterations"] for s in fp64_stats])) if fp64_stats else None
    value = args.steps / total
    fp64_value = len(fp64_times) / sum(fp64_times) if fp64_times else None

    # dominant kernel: the first FP32 stencil pass k_matvec_dots4<BS, 0>
    # (V = A*(P/A0) and <Rtilde,V>, four cells per thread): 6 off-diagonal planes + row excess
    # + operand + dot partner + result = 10*4 = 40 B/cell (start and end events
    # recorded by each launch's dispatch in this run; A0 = -1 on every
    # unit-scaled row of this system, so 1/A0 is not read).
    # Only launches that worked count: the ones issued after a solve was done
    # return at once and are reported apart (idle_launches).
    kern_avg_s = prof["total_ms"] * 1e-3 / max(1, prof["launches"])
    kern_bytes = prof["alg_bytes"] / max(1, prof["launches"])
    achieved = kern_bytes / kern_avg_s if kern_avg_s > 0 else 0.0
    traffic = None
    tf = ROOT / "profiles" / "stencil_fp32_256cubed_traffic.json"
    if tf.exists():
        # measured HBM bytes per launch of the shape this run used (autotuned
        # shapes differ in their halo re-reads), else of the counter runs' shape
        try:
            tj = json.loads(tf.read_text())
            cells = shapes.get("mixed", "256x1").split("x")[0]
            traffic = tj.get("bytes_per_launch_by_shape", {}).get(cells, tj.get("bytes_per_launch"))
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
                               int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1))
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
            "launch_shape": shapes or "default",
            "latency_ms_p50": 1e3 * _percentile(times, 50),
            "latency_ms_p99": 1e3 * _percentile(times, 99),
            "fp64_solves_per_s": fp64_value,
            "fp64_iterations": fp64_its,
            "speedup_vs_fp64": (value / fp64_value) if fp64_value else None,
            "solve_alg_GBps": solve_rate / 1e9,
            "solve_roofline_frac": solve_rate / HBM_PEAK,
            "true_rel_residual": float(np.max([s["true_rnorm"] / s["rnorm0"] for s in stats])),
            "max_rel_error_vs_fp64": err_vs_fp64,
            "generation_s": gen_s,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_matvec_dots4<BS, 0> (FP32 stencil V = A*(P/A0) + <Rtilde,V>)",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic,
            "launches": prof["launches"],
            "idle_launches": prof["idle_launches"],
            "avg_launch_us": kern_avg_s * 1e6,
            "alg_bytes_per_launch": kern_bytes,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
