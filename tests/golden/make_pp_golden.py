"""Write golden hashes of the reference's gas pressure-correction assembly.

Runs oracle/_ref/leq_ref_harness op 10 (INIT_AB_M, CONV_PP_G, SOURCE_PP_G of
model/init_ab_m.f, model/conv_pp_g.f, model/source_pp_g.f, with the
neighbour tables of the reference's SET_INDEX1A on each case's FLAG) on the
seeded cases of tests/pp_cases.py and stores the SHA-256 of the assembled
A_M, B_M and of B_MMAX on the fluid cells in tests/golden/pp_ref_golden.npz
(the outputs are compared bit for bit).  Run where the MFIX tree is mounted:
    python tests/golden/make_pp_golden.py
"""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from tests.pp_cases import pp_cases, pp_input_sha  # noqa: E402

OUT = Path(__file__).resolve().parent / "pp_ref_golden.npz"


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main() -> None:
    if oracle.build_ref() is None:
        raise SystemExit("reference tree not present: golden vectors cannot be regenerated here")
    data: dict[str, np.ndarray] = {}
    names = []
    for case in pp_cases():
        g = case["grid"]
        n3 = oracle.n3_of(g)
        r = oracle.run_reference("assemble_pp", g, np.zeros(7 * n3), np.zeros(n3), np.zeros(n3),
                                 pp=dict(flag=case["flag"], odt=case["odt"], ijk_p_g=case["ijk_p_g"],
                                         fields=case["fields"]))
        fl = case["flag"] == 1
        key = case["name"]
        data[f"{key}/in_sha"] = np.asarray(pp_input_sha(case))
        data[f"{key}/a_sha"] = np.asarray(sha(r["a"]))
        data[f"{key}/b_sha"] = np.asarray(sha(r["b"]))
        data[f"{key}/bmmax_sha"] = np.asarray(sha(r["x"][fl]))
        data[f"{key}/a0_sum"] = np.float64(r["a"][3 * n3:4 * n3].sum())
        names.append(key)
    data["cases"] = np.array(names)
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({OUT.stat().st_size / 1e3:.1f} kB, {len(names)} cases)")


if __name__ == "__main__":
    main()
