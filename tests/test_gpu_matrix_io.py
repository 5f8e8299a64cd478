"""Replaying MFIX matrix dumps (matrixExtract CSV, WRITE_AB_M log) through
the device solver: the replayed system solves to the sparse direct solution
of the same parsed arrays."""
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from mfix_amd import matrix_io
from tests.cases import to_csr

pytestmark = pytest.mark.gpu

DUMPS = Path(__file__).resolve().parent / "golden" / "dumps"
CASES = [("p3d_6x5x4", 6, 5, 4, False, "DIAG"), ("m3d_7x6x5", 7, 6, 5, False, "LINE"),
         ("m2d_10x8", 10, 8, 1, True, "LINE")]


@pytest.mark.parametrize("name,im,jm,km,no_k,pc", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("mixed", [False, True])
def test_replay_matrix_extract(gpu, name, im, jm, km, no_k, pc, mixed):
    d = DUMPS / name
    a, b = matrix_io.read_matrix_extract(d / "Am_HARNESS.csv", d / "Bm_HARNESS.csv", im, jm, km, no_k=no_k)
    r = matrix_io.replay(a, b, im, jm, km, no_k=no_k, pc=pc, tol=1e-10, itmax=500, mixed=mixed)
    assert r["ier"] == 0, r["stats"]
    x_ref = spla.spsolve(to_csr(a, im, jm, km, no_k=no_k).tocsc(), b)
    assert np.abs(r["x"] - x_ref).max() <= 1e-7 * np.abs(x_ref).max()


def test_replay_write_ab_m_cli(gpu, capsys):
    log = DUMPS / "p3d_6x5x4" / "write_ab_m.txt"
    assert matrix_io.main(["--log", str(log), "--pc", "DIAG"]) == 0
    out = capsys.readouterr().out
    assert '"ier": 0' in out
