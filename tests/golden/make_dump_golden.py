"""Write MFIX matrix dumps with the reference's own writers.

Runs oracle/_ref/leq_ref_harness op 12 (initExtract over the whole domain +
matrixExtract, model/dbg_mod.f) and op 13 (WRITE_AB_M, model/write_ab_m.f)
on three seeded systems of tests/cases.py (a 3-D pressure system, a 3-D
upwind momentum system, a 2-D momentum system) and stores the text files the
reference wrote under tests/golden/dumps/<case>/ (Am_HARNESS.csv,
Bm_HARNESS.csv and, for WRITE_AB_M's log output, write_ab_m.txt), with the systems
themselves in tests/golden/dumps/sources.npz, so that mfix_amd.matrix_io can
be checked against the real formats.  Run where the MFIX tree is mounted:
    python tests/golden/make_dump_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from tests.cases import make_system  # noqa: E402

OUT = Path(__file__).resolve().parent / "dumps"
# name, kind, imax, jmax, kmax, no_k, seed
CASES = [("p3d_6x5x4", 0, 6, 5, 4, False, 31), ("m3d_7x6x5", 1, 7, 6, 5, False, 32),
         ("m2d_10x8", 1, 10, 8, 1, True, 33)]


def main() -> None:
    if oracle.build_ref() is None:
        raise SystemExit("reference tree not present: the dumps cannot be regenerated here")
    sources = {}
    for name, kind, im, jm, km, no_k, seed in CASES:
        g = oracle.grid(im, jm, km, no_k=no_k)
        n3 = oracle.n3_of(g)
        a, b = make_system(kind, im, jm, km, no_k=no_k, seed=seed)
        sources[f"{name}/a"] = a
        sources[f"{name}/b"] = b
        d = OUT / name
        d.mkdir(parents=True, exist_ok=True)
        for op in ("extract", "write_ab_m"):
            r = oracle.run_reference(op, g, a, b, np.zeros(n3))
            for fname, text in r["files"].items():
                # the log MFIX would write to is kept as write_ab_m.txt
                (d / ("write_ab_m.txt" if fname.endswith(".log") else fname)).write_text(text)
    np.savez_compressed(OUT / "sources.npz", **sources)
    total = sum(p.stat().st_size for p in OUT.rglob("*") if p.is_file())
    print(f"wrote {OUT} ({total / 1e3:.1f} kB)")


if __name__ == "__main__":
    main()
