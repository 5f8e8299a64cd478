"""Reference LEQ_BICGS results on the benchmark configurations, and the
CPU-baseline fairness ratio.

Runs the reference's own LEQ_BICGS (oracle/_ref/leq_ref_harness: the MFIX
closure compiled from the reference tree with flang -O2 -fopenmp, driven by
oracle/ref_harness.f90, op 1) on the systems of configurations 0, 1 and 2,
and the oracle (oracle/leq_oracle.c, gcc -O2 -fopenmp) on the same inputs
with the same R~ noise (oracle.flang_noise regenerates the array the
reference drew), on the same cores.  Writes

* tests/golden/config_ref_golden.npz: per configuration IER, ITER_TOT,
  Rnorm0 and the true relative residual of the reference's solution on the
  scaled system, a strided sample of its solution (every STRIDE-th cell) and
  checksums of the whole solution, plus the noise seed;
* tests/golden/cpu_fairness.json: wall time of the reference and of the
  oracle on the same cores (full solves of configurations 0 and 1, median of
  REPEATS; configuration 2 per iteration, from the difference of two ITMAX
  runs), and their ratio.

The fixture pins what only the compiled reference can pin at the
configurations' real sizes (iteration counts and IER, model/leq_bicgs.f:
440-495); the systems themselves are regenerated from their seeds.  Run where
the MFIX tree is mounted (about 15 minutes on 8 cores):

    OMP_NUM_THREADS=8 python tests/golden/make_config_golden.py
"""
from __future__ import annotations

import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from mfix_amd import systems  # noqa: E402

OUT = Path(__file__).resolve().parent / "config_ref_golden.npz"
FAIR = Path(__file__).resolve().parent / "cpu_fairness.json"
STRIDE = 997
REPEATS = 5
# configuration -> (PC, CMETHOD, ITMAX, noise seed)
CONFIGS = {0: ("DIAG", "RSRS", 500, 17), 1: ("LINE", "RSRS", 500, 1017), 2: ("DIAG", "RSRS", 5000, 2017)}


def _true_rel(g, a_s, b_s, x, rn0):
    r = b_s - oracle.matvec(g, a_s, x)
    return float(np.sqrt(np.dot(r, r)) / rn0)


def main() -> None:
    if oracle.build_ref() is None:
        raise SystemExit("reference tree not present: the fixture cannot be regenerated here")
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    oracle.set_threads(threads)
    data: dict[str, np.ndarray] = {}
    fair = {"threads": threads, "cpu": platform.processor() or platform.machine(), "configs": {}}
    try:
        fair["cpu"] = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    for cfg, (pc, cm, itmax, seed) in CONFIGS.items():
        n = {0: 32, 1: 128, 2: 256}[cfg]
        a, b = systems.config_system(cfg)
        g = oracle.grid(n, n, n)
        x0 = np.zeros(b.size)
        t = time.time()
        ref = oracle.run_reference("bicgs", g, a, b, x0, pc=pc, cmethod=cm, tol=1e-10, itmax=itmax, seed=seed)
        print(f"config {cfg}: reference IER {ref['ier']} ITER_TOT {ref['iter_tot']} "
              f"({ref['dot'][1]:.2f} s in the call, {time.time() - t:.1f} s overall)", flush=True)
        noise = oracle.flang_noise(b.size, seed)
        assert np.array_equal(noise, ref["noise"]), "flang_noise does not reproduce the harness noise"
        a_s, b_s = ref["a"], ref["b"]
        rn0 = float(np.sqrt(np.dot(b_s, b_s)))  # x0 = 0: R0 = B of the scaled system
        x = ref["x"]
        idx = np.arange(0, b.size, STRIDE)
        key = f"config{cfg}"
        data.update({
            f"{key}/ier": np.int32(ref["ier"]), f"{key}/iter_tot": np.int32(ref["iter_tot"]),
            f"{key}/rnorm0": np.float64(rn0), f"{key}/true_rel": np.float64(_true_rel(g, a_s, b_s, x, rn0)),
            f"{key}/idx": idx.astype(np.int64), f"{key}/x_sample": x[idx],
            f"{key}/x_sum": np.float64(x.sum()), f"{key}/x_abs_sum": np.float64(np.abs(x).sum()),
            f"{key}/x_max": np.float64(np.abs(x).max()), f"{key}/noise_seed": np.int32(seed),
            f"{key}/noise_head": noise[:8].copy(), f"{key}/itmax": np.int32(itmax),
            f"{key}/pc": np.array(pc), f"{key}/cmethod": np.array(cm),
            f"{key}/ref_seconds": np.float64(ref["dot"][1]),
        })
        # the oracle on the same inputs and noise: same IER and iterations,
        # solution to rounding (its dot products sum in the same order)
        xo, ao, bo, res = oracle.leq_bicgs(g, pc, cm, 1e-10, itmax, x0, a, b, noise)
        assert np.array_equal(ao, a_s) and np.array_equal(bo, b_s)
        assert res["ier"] == ref["ier"], (res, ref["ier"])
        it_o = res["iter_tot_increment"] if res["ier"] == 0 else res["iterations"]
        diff = float(np.max(np.abs(xo - x)) / np.max(np.abs(x)))
        print(f"config {cfg}: oracle IER {res['ier']} ITER_TOT {res['iter_tot_increment']} "
              f"its {res['iterations']}, max-norm difference {diff:.3e}, {res['seconds_total']:.2f} s", flush=True)
        data[f"{key}/oracle_iterations"] = np.int32(res["iterations"])
        data[f"{key}/oracle_max_rel_diff"] = np.float64(diff)
        entry = {"pc": pc, "cmethod": cm, "n3": int(b.size), "iterations": int(res["iterations"]),
                 "reference_first_s": float(ref["dot"][1]), "oracle_first_s": float(res["seconds_total"])}
        if cfg in (0, 1):
            rt, ot = [], []
            for _ in range(REPEATS):
                rt.append(oracle.run_reference("bicgs", g, a, b, x0, pc=pc, cmethod=cm, tol=1e-10, itmax=itmax,
                                               seed=seed)["dot"][1])
                ot.append(oracle.leq_bicgs(g, pc, cm, 1e-10, itmax, x0, a, b, noise)[3]["seconds_total"])
            entry.update(reference_median_s=float(np.median(rt)), oracle_median_s=float(np.median(ot)),
                         repeats=REPEATS)
        else:
            # per iteration: difference of two ITMAX runs (the set-up, which
            # includes the serial random_number draw, cancels)
            lo, hi = 10, 60
            rl = [oracle.run_reference("bicgs", g, a, b, x0, pc=pc, cmethod=cm, tol=1e-10, itmax=m,
                                       seed=seed)["dot"][1] for m in (lo, hi)]
            ol = [oracle.leq_bicgs(g, pc, cm, 1e-10, m, x0, a, b, noise)[3]["seconds_total"] for m in (lo, hi)]
            entry.update(reference_ms_per_iteration=1e3 * (rl[1] - rl[0]) / (hi - lo),
                         oracle_ms_per_iteration=1e3 * (ol[1] - ol[0]) / (hi - lo),
                         reference_setup_s=rl[0] - lo * (rl[1] - rl[0]) / (hi - lo),
                         oracle_setup_s=ol[0] - lo * (ol[1] - ol[0]) / (hi - lo), sample_itmax=[lo, hi])
        if "reference_median_s" in entry:
            entry["ratio_reference_over_oracle"] = entry["reference_median_s"] / entry["oracle_median_s"]
        else:
            entry["ratio_reference_over_oracle"] = (entry["reference_ms_per_iteration"] /
                                                    entry["oracle_ms_per_iteration"])
        fair["configs"][str(cfg)] = entry
        print(json.dumps(entry), flush=True)
    np.savez_compressed(OUT, **data)
    FAIR.write_text(json.dumps(fair, indent=1) + "\n")
    print(f"wrote {OUT} and {FAIR}")


if __name__ == "__main__":
    main()
