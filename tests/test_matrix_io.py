"""Readers of MFIX's matrix dumps (mfix_amd/matrix_io.py) against files the
reference's own writers produced: matrixExtract (model/dbg_mod.f:228-552)
and WRITE_AB_M (model/write_ab_m.f:20-145), run by the compiled reference
harness on seeded systems (tests/golden/make_dump_golden.py).  The readers
must place every value in its cell and plane: the result equals the source
system to the text precision of each format (e14.6: six significant digits;
G9.2: two), zeros stay exactly zero, and WRITE_AB_M's grid is inferred."""
from pathlib import Path

import numpy as np
import pytest

from mfix_amd import matrix_io

DUMPS = Path(__file__).resolve().parent / "golden" / "dumps"
CASES = [("p3d_6x5x4", 6, 5, 4, False), ("m3d_7x6x5", 7, 6, 5, False), ("m2d_10x8", 10, 8, 1, True)]


@pytest.fixture(scope="module")
def sources():
    return np.load(DUMPS / "sources.npz", allow_pickle=False)


def _close(got, want, rel):
    assert got.shape == want.shape
    np.testing.assert_array_equal(got == 0, want == 0)
    nz = want != 0
    err = np.abs(got[nz] - want[nz]) / np.abs(want[nz])
    assert err.max() <= rel, err.max()


@pytest.mark.parametrize("name,im,jm,km,no_k", CASES, ids=[c[0] for c in CASES])
def test_matrix_extract_reader(sources, name, im, jm, km, no_k):
    d = DUMPS / name
    a, b = matrix_io.read_matrix_extract(d / "Am_HARNESS.csv", d / "Bm_HARNESS.csv", im, jm, km, no_k=no_k)
    _close(a, sources[f"{name}/a"], 5.0e-6)
    _close(b, sources[f"{name}/b"], 5.0e-6)


@pytest.mark.parametrize("name,im,jm,km,no_k", CASES, ids=[c[0] for c in CASES])
def test_write_ab_m_reader(sources, name, im, jm, km, no_k):
    a, b, dims = matrix_io.read_write_ab_m(DUMPS / name / "write_ab_m.txt")
    assert dims == (im, jm, 0 if no_k else km, no_k)
    _close(a, sources[f"{name}/a"], 5.0e-2)
    _close(b, sources[f"{name}/b"], 5.0e-2)


def test_matrix_extract_rejects_wrong_shapes():
    d = DUMPS / "p3d_6x5x4"
    with pytest.raises(ValueError):
        matrix_io.read_matrix_extract(d / "Am_HARNESS.csv", d / "Bm_HARNESS.csv", 6, 5, 5)     # wrong KMAX
    with pytest.raises(ValueError):
        matrix_io.read_matrix_extract(d / "Am_HARNESS.csv", d / "Bm_HARNESS.csv", 6, 5, 4, no_k=True)
    d2 = DUMPS / "m2d_10x8"
    with pytest.raises(ValueError):
        matrix_io.read_matrix_extract(d2 / "Am_HARNESS.csv", d2 / "Bm_HARNESS.csv", 10, 8, 1, no_k=False)


def test_matrix_extract_subregion(tmp_path, sources):
    """Placement of a dump of a sub-box (initExtract with bounds): its rows
    are the box's cells in K, I, J order and cells outside it stay zero.  The
    file is cut from the whole-domain dump, so its coefficients toward cells
    outside the box are not zeroed as matrixExtract would zero them; the
    reader places whatever values the rows hold."""
    name, im, jm, km = "p3d_6x5x4", 6, 5, 4
    d = DUMPS / name
    ni, nj, nk = im + 2, jm + 2, km + 2
    region = (2, 5, 3, 6, 2, 4)
    am = np.loadtxt(d / "Am_HARNESS.csv", delimiter=",").reshape(nk, ni, nj, 7)
    bm = np.loadtxt(d / "Bm_HARNESS.csv").reshape(nk, ni, nj)
    ilb, iub, jlb, jub, klb, kub = region
    box = (slice(klb - 1, kub), slice(ilb - 1, iub), slice(jlb - 1, jub))
    np.savetxt(tmp_path / "Am.csv", am[box].reshape(-1, 7), delimiter=",", fmt="%14.6E")
    np.savetxt(tmp_path / "Bm.csv", bm[box].reshape(-1), fmt="%14.6E")
    a, b = matrix_io.read_matrix_extract(tmp_path / "Am.csv", tmp_path / "Bm.csv", im, jm, km, region=region)
    A = a.reshape(7, nk, ni, nj)
    want = np.moveaxis(am, 3, 0)
    np.testing.assert_array_equal(A[(slice(None),) + box], want[(slice(None),) + box])
    mask = np.ones((nk, ni, nj), bool)
    mask[box] = False
    assert (A[:, mask] == 0).all() and (b.reshape(nk, ni, nj)[mask] == 0).all()
