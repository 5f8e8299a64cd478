"""The oracle's gas pressure-correction assembly (INIT_AB_M, CONV_PP_G,
SOURCE_PP_G; model/init_ab_m.f, model/conv_pp_g.f, model/source_pp_g.f)
against the compiled reference's outputs, bit for bit
(tests/golden/pp_ref_golden.npz, written by tests/golden/make_pp_golden.py),
and against the equation's invariants."""
import hashlib
from pathlib import Path

import numpy as np
import pytest

import oracle
from tests.pp_cases import pp_cases, pp_input_sha

GOLDEN = Path(__file__).resolve().parent / "golden" / "pp_ref_golden.npz"
CASES = pp_cases()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_assembly_matches_reference_golden(golden, case):
    key = case["name"]
    assert str(golden[f"{key}/in_sha"]) == pp_input_sha(case)
    a, b, bm, bad = oracle.assemble_pp_g(case["grid"], case["fields"], case["flag"], case["odt"], case["ijk_p_g"])
    fl = case["flag"] == 1
    assert sha(a) == str(golden[f"{key}/a_sha"])
    assert sha(b) == str(golden[f"{key}/b_sha"])
    assert sha(bm[fl]) == str(golden[f"{key}/bmmax_sha"])
    assert bad == 0


@pytest.mark.parametrize("case", CASES[:4], ids=[c["name"] for c in CASES[:4]])
def test_assembly_invariants(case):
    """Off-diagonals >= 0; each fluid row's A0 is minus the sum of its
    off-diagonals before the pressure-flow faces are cut; non-fluid rows are
    -1/0; the pressure-correction operator is symmetric between two fluid
    cells that both keep the shared face."""
    g = case["grid"]
    n3 = oracle.n3_of(g)
    a, b, bm, _ = oracle.assemble_pp_g(g, case["fields"], case["flag"], case["odt"], case["ijk_p_g"])
    A = a.reshape(7, n3)
    fl = case["flag"] == 1
    assert (np.delete(A, 3, axis=0) >= 0).all()
    assert (A[3][~fl] == -1).all() and (b[~fl] == 0).all()
    assert (A[:, ~fl][[0, 1, 2, 4, 5, 6]] == 0).all()
    nj = g.jmax + 2
    # east/west pair of two interior fluid cells (no cyclic wrap between them)
    cells = np.flatnonzero(fl)
    checked = 0
    for c in cells:
        e = c + nj
        if e < n3 and fl[e] and c != case["ijk_p_g"] - 1 and e != case["ijk_p_g"] - 1:
            assert A[2, e] == A[4, c]
            checked += 1
    assert checked > 0
