"""Write golden hashes of the reference's scalar-transport assembly.

Runs oracle/_ref/leq_ref_harness op 11 (INIT_AB_M + CONV_DIF_PHI with
DISC = 0, SOURCE_PHI and UNDER_RELAX_S of
model/init_ab_m.f, model/conv_dif_phi.f, model/source_phi.f and
model/under_relax.f, with the neighbour tables of the reference's
SET_INDEX1A on each case's FLAG) on the seeded cases of tests/phi_cases.py,
chained as tests/phi_cases.reference_chain describes, and stores the
SHA-256 of A_M and B_M after each stage in tests/golden/phi_ref_golden.npz
(the outputs are compared bit for bit).  Run where the MFIX tree is mounted:
    python tests/golden/make_phi_golden.py
"""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from tests.phi_cases import phi_cases, phi_input_sha, reference_chain  # noqa: E402

OUT = Path(__file__).resolve().parent / "phi_ref_golden.npz"


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main() -> None:
    if oracle.build_ref() is None:
        raise SystemExit("reference tree not present: golden vectors cannot be regenerated here")
    data: dict[str, np.ndarray] = {}
    names = []
    for case in phi_cases():
        g = case["grid"]
        n3 = oracle.n3_of(g)

        def run(stage, a, b):
            a = np.zeros(7 * n3) if a is None else a
            b = np.zeros(n3) if b is None else b
            r = oracle.run_reference("phi", g, a, b, np.zeros(n3),
                                     phi=dict(flag=case["flag"], stage=stage, ur=case["ur"], fields=case["fields"]))
            return r["a"], r["b"]

        key = case["name"]
        data[f"{key}/in_sha"] = np.asarray(phi_input_sha(case))
        for s, (a, b) in enumerate(reference_chain(case, run), start=1):
            data[f"{key}/a{s}_sha"] = np.asarray(sha(a))
            data[f"{key}/b{s}_sha"] = np.asarray(sha(b))
            data[f"{key}/a{s}_sum"] = np.float64(np.abs(a).sum())
        names.append(key)
    data["cases"] = np.array(names)
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({OUT.stat().st_size / 1e3:.1f} kB, {len(names)} cases)")


if __name__ == "__main__":
    main()
