"""GPU gas pressure-correction equation (INIT_AB_M + CONV_PP_G +
SOURCE_PP_G, and SOLVE_PP_G around it; model/conv_pp_g.f,
model/source_pp_g.f, model/solve_pp_g.f) against the oracle, through the C
ABI.  The assembly is element-wise arithmetic in the reference's order: A_M,
B_M and B_MMAX must match bit for bit.  SOLVE_PP_G's residual RESG is a sum
(relative 1e-13) and its solve meets TOL on the scaled system."""
import numpy as np
import pytest

import oracle
from tests.pp_cases import pp_cases, pp_fields, pp_flags

pytestmark = pytest.mark.gpu

CASES = pp_cases()


def _upload(gpu, case):
    fields = {k: gpu.DeviceArray.from_host(np.ascontiguousarray(v)) for k, v in case["fields"].items()}
    flag = gpu.DeviceArray.from_host(np.ascontiguousarray(case["flag"], dtype=np.int32))
    return fields, flag


def _grid_init(gpu, g):
    gpu.grid_init(g.imax, g.jmax, g.kmax, no_k=bool(g.no_k),
                  cyclic=(bool(g.cyclic_x), bool(g.cyclic_y), bool(g.cyclic_z)))


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_assembly_bit_exact(gpu, case):
    g = case["grid"]
    n3 = oracle.n3_of(g)
    _grid_init(gpu, g)
    fields, flag = _upload(gpu, case)
    da, db = gpu.DeviceArray(7 * n3), gpu.DeviceArray(n3)
    dbm = gpu.DeviceArray.from_host(np.full(n3, -7.0))
    nbad = gpu.assemble_pp_g_dev(fields, flag, case["odt"], case["ijk_p_g"], da, db, dbm)
    a, b, bm, bad = oracle.assemble_pp_g(g, case["fields"], case["flag"], case["odt"], case["ijk_p_g"])
    fl = case["flag"] == 1
    np.testing.assert_array_equal(da.download(), a)
    np.testing.assert_array_equal(db.download(), b)
    got_bm = dbm.download()
    np.testing.assert_array_equal(got_bm[fl], bm[fl])
    assert (got_bm[~fl] == -7.0).all()          # B_MMAX is written on fluid cells only
    assert nbad == bad


def _solve_case(seed=3):
    im, jm, km = 24, 20, 22
    g = oracle.grid(im, jm, km)
    n3 = oracle.n3_of(g)
    flag = pp_flags(im, jm, km, False, 600 + seed)
    fields = pp_fields(n3, 650 + seed)
    return g, flag, fields


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("normg", [0.0, 0.5])
def test_solve_pp_g(gpu, mixed, normg):
    g, flag_h, fields_h = _solve_case()
    n3 = oracle.n3_of(g)
    _grid_init(gpu, g)
    case = dict(fields=fields_h, flag=flag_h)
    fields, flag = _upload(gpu, case)
    dpp = gpu.DeviceArray.from_host(np.full(n3, 3.0))   # SOLVE_PP_G resets PP_G to 0
    da, db, dbm = gpu.DeviceArray(7 * n3), gpu.DeviceArray(n3), gpu.DeviceArray(n3)
    out = gpu.solve_pp_g_dev(fields, flag, 100.0, 0, normg, dpp, da, db, dbm, tol=1e-10, pc="DIAG", mixed=mixed)
    # oracle: assembly, the two CALC_RESID_PP calls (solve_pp_g.f:100-112)
    a, b, bm, _ = oracle.assemble_pp_g(g, fields_h, flag_h, 100.0, 0)
    norm = normg
    if normg == 0.0:
        norm = oracle.calc_resid(g, "PP", np.zeros(n3), a, bm, flag_h, norm=1.0)["resid"] / 10.0
    resg = oracle.calc_resid(g, "PP", np.zeros(n3), a, b, flag_h, norm=norm)["resid"]
    np.testing.assert_allclose(out["resg"], resg, rtol=1e-13)
    assert out["ier"] == 0, out
    a_s, b_s, _ = oracle.unit_scale(g, a, b)
    np.testing.assert_array_equal(da.download(), a_s)
    np.testing.assert_array_equal(db.download(), b_s)
    x = dpp.download()
    r = b_s - oracle.matvec(g, a_s, x)
    assert np.sqrt(r @ r) <= 1e-10 * np.sqrt(b_s @ b_s) * 1.0000001
    assert (x[flag_h != 1] == 0.0).all()


def test_solve_pp_g_rejects_other_methods(gpu):
    g, flag_h, fields_h = _solve_case()
    n3 = oracle.n3_of(g)
    _grid_init(gpu, g)
    fields, flag = _upload(gpu, dict(fields=fields_h, flag=flag_h))
    dpp, da, db, dbm = (gpu.DeviceArray(n) for n in (n3, 7 * n3, n3, n3))
    with pytest.raises(Exception):
        gpu.solve_pp_g_dev(fields, flag, 100.0, 0, 0.5, dpp, da, db, dbm, method=1)
