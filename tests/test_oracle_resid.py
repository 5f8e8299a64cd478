"""The oracle's CALC_RESID_C / _S / _PP (model/calc_resid.f) against the
compiled reference's own outputs (tests/golden/resid_ref_golden.npz, written
by tests/golden/make_resid_golden.py through oracle/ref_harness.f90), an
independent numpy formulation, and the routine's closed-form branches."""
from pathlib import Path

import numpy as np
import pytest

import oracle
from tests.cases import GRIDS, make_system
from tests.resid_cases import flags_for, input_sha, resid_cases

GOLDEN = Path(__file__).resolve().parent / "golden" / "resid_ref_golden.npz"


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


CASES = resid_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_matches_reference_golden(golden, case):
    key = case["name"]
    assert str(golden[f"{key}/in_sha"]) == input_sha(case), "case generator drifted from the fixture"
    o = oracle.calc_resid(case["grid"], case["kind"], case["x"], case["a"], case["b"], case["flag"], tol=case["tol"],
                          ep_s=case["ep_s"] if case["m"] > 0 else None, norm=case["norm"], debug=case["debug"])
    want = golden[f"{key}/out"]
    got = np.array([o["num"], o["den"], o["resid"], o["max_resid"]])
    np.testing.assert_array_equal(got, want)
    assert o["ijk_resid"] == int(golden[f"{key}/ijk"])


def _numpy_resid_c(im, jm, km, a, b, x, flag):
    """CALC_RESID_C on a closed box (ghost walls, fluid interior, no
    cyclic axes), written with array shifts: a wall neighbour is replaced by
    the cell itself (model/set_index1a.f:93-97)."""
    ni, nj, nk = im + 2, jm + 2, km + 2
    A = a.reshape(7, nk, ni, nj)
    X = x.reshape(nk, ni, nj)
    B = b.reshape(nk, ni, nj)
    F = flag.reshape(nk, ni, nj)
    s = (slice(1, nk - 1), slice(1, ni - 1), slice(1, nj - 1))
    xc = X[s]

    def nbv(dk, di, dj):
        sh = (slice(1 + dk, nk - 1 + dk), slice(1 + di, ni - 1 + di), slice(1 + dj, nj - 1 + dj))
        return np.where(F[sh] >= 100, xc, X[sh])

    ax = A[3][s] * xc
    sm = ax + A[4][s] * nbv(0, 1, 0) + A[2][s] * nbv(0, -1, 0) + A[5][s] * nbv(0, 0, 1) + A[1][s] * nbv(0, 0, -1)
    r = B[s] - sm - (A[6][s] * nbv(1, 0, 0) + A[0][s] * nbv(-1, 0, 0))
    fl = F[s] == 1
    return np.abs(r)[fl], np.abs(ax)[fl]


@pytest.mark.parametrize("gi", [0, 4])
def test_resid_c_matches_numpy_shifts(gi):
    name, kind, im, jm, km, no_k, cyc = GRIDS[gi]
    a, b = make_system(kind, im, jm, km, no_k)
    x = np.random.default_rng(3).uniform(-1, 1, b.size)
    flag = flags_for(im, jm, km, no_k, 9, pattern="fluid")
    # interior walls too (no flow cells: the numpy form has no true-corner rule)
    f3 = flag.reshape(km + 2, im + 2, jm + 2)
    f3[2, 2, 3] = 100
    f3[km, im, 1 + jm // 2] = 101
    g = oracle.grid(im, jm, km)
    o = oracle.calc_resid(g, "C", x, a, b, flag, debug=True)
    num1, den1 = _numpy_resid_c(im, jm, km, a, b, x, flag)
    assert o["ncells"] == num1.size
    np.testing.assert_allclose(o["num"], num1.sum(), rtol=1e-13)
    np.testing.assert_allclose(o["den"], den1.sum(), rtol=1e-13)
    np.testing.assert_allclose(o["max_resid"], num1.size * num1.max() / den1.sum(), rtol=1e-13)


def test_pp_closed_form():
    name, kind, im, jm, km, no_k, cyc = GRIDS[0]
    a, b = make_system(kind, im, jm, km, no_k)
    flag = flags_for(im, jm, km, no_k, 1, pattern="fluid")
    g = oracle.grid(im, jm, km)
    fl = flag == 1
    o = oracle.calc_resid(g, "PP", np.zeros(b.size), a, b, flag, norm=2.0, debug=False)
    assert o["ncells"] == fl.sum()
    np.testing.assert_allclose(o["num"], np.abs(b[fl]).sum(), rtol=1e-14)
    np.testing.assert_allclose(o["resid"], np.abs(b[fl]).sum() / (fl.sum() * 2.0), rtol=1e-14)
    first = int(np.flatnonzero(fl & (np.abs(b) == np.abs(b[fl]).max()))[0])
    assert o["ijk_resid"] == first + 1 and o["max_resid"] == np.abs(b[fl]).max()


def test_no_fluid_cells_give_zero_residual():
    name, kind, im, jm, km, no_k, cyc = GRIDS[0]
    a, b = make_system(kind, im, jm, km, no_k)
    flag = flags_for(im, jm, km, no_k, 1, pattern="allwall")
    g = oracle.grid(im, jm, km)
    for kd in ("C", "S"):
        o = oracle.calc_resid(g, kd, np.ones(b.size), a, b, flag, debug=True)
        assert (o["num"], o["den"], o["resid"], o["max_resid"], o["ijk_resid"], o["ncells"]) == (0, 0, 0, 0, 0, 0)
