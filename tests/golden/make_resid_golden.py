"""Write golden vectors of the reference's own residual routines.

Runs oracle/_ref/leq_ref_harness (ops 7-9: CALC_RESID_C, CALC_RESID_S,
CALC_RESID_PP of model/calc_resid.f, with the neighbour tables built by the
reference's SET_INDEX1A on each case's FLAG array) and stores its outputs
in tests/golden/resid_ref_golden.npz.  The inputs are the seeded cases of
tests/resid_cases.py (the SHA-256 of each case's A_M, B_M, VAR, FLAG and
EP_s is stored to pin the generator).  Run where the MFIX tree is mounted:
    python tests/golden/make_resid_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from tests.cases import GRIDS  # noqa: E402
from tests.resid_cases import input_sha, resid_cases  # noqa: E402

OUT = Path(__file__).resolve().parent / "resid_ref_golden.npz"


def main() -> None:
    if oracle.build_ref() is None:
        raise SystemExit("reference tree not present: golden vectors cannot be regenerated here")
    data: dict[str, np.ndarray] = {}
    names = []
    for case in resid_cases():
        key = case["name"]
        g = case["grid"]
        op = {"C": "resid_c", "S": "resid_s", "PP": "resid_pp"}[case["kind"]]
        r = oracle.run_reference(op, g, case["a"], case["b"], case["x"],
                                 resid=dict(flag=case["flag"], debug=case["debug"], m=case["m"], tol=case["tol"],
                                            norm=case["norm"], ep_s=case["ep_s"]))
        data[f"{key}/in_sha"] = np.asarray(input_sha(case))
        data[f"{key}/out"] = np.array([r["dot"][0], r["dot"][1], r["x"][0], r["x"][1]])
        data[f"{key}/ijk"] = np.int32(r["ier"])
        names.append(key)
    data["cases"] = np.array(names)
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({OUT.stat().st_size / 1e3:.1f} kB, {len(names)} cases)")


if __name__ == "__main__":
    main()
