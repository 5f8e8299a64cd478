"""The Fortran drop-in (mfix_amd/fortran/leq_bicgs_shim.f90) on the GPU.

SOLVE_LIN_EQ calls LEQ_BICGS with an implicit interface, so the linker sees
`leq_bicgs_` with the CHARACTER lengths appended as hidden trailing
arguments (model/solve_lin_eq.f:151-152).  This test calls that symbol the
way compiled MFIX code does and checks that it returns exactly what the C
entry mfix_leq_bicgs returns for the same call: VAR, the row-scaled A_M/B_M,
IER and ITER_TOT.
"""
import ctypes as C

import numpy as np
import pytest

from mfix_amd import FORTRAN_LIB_PATH, systems

pytestmark = pytest.mark.gpu


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(v):
    return C.byref(C.c_int(v))


@pytest.mark.parametrize("pc,kind", [("LINE", 1), ("DIAG", 0)])
def test_fortran_leq_bicgs_matches_c_entry(gpu, pc, kind):
    if not FORTRAN_LIB_PATH.exists():
        pytest.skip("the flang-built shim is not present")
    f = C.CDLL(str(FORTRAN_LIB_PATH))
    im, jm, km = 10, 9, 8
    ierr = C.c_int(-99)
    false, true = 0, 1
    # MFIX_LEQ_GRID_INIT_F(IMAX,JMAX,KMAX,NO_K,CYCLIC_X,CYCLIC_Y,CYCLIC_Z,IS_SERIAL,MINIMIZE,ICHECK,IERR)
    f.mfix_leq_grid_init_f_(_i(im), _i(jm), _i(km), _i(false), _i(false), _i(false), _i(false), _i(true),
                            _i(false), _i(1), C.byref(ierr))
    assert ierr.value == 0
    a, b = systems.gen_system(kind, im, jm, km, 77)
    var, am, bm = np.zeros(b.size), a.copy(), b.copy()
    gpu.reset_iter_tot()
    ier = C.c_int(5)
    vname, cmethod = b"U_g", b"RSRS"
    f.leq_bicgs_(C.c_char_p(vname), _i(2), _dp(var), _dp(am), _dp(bm), C.c_char_p(cmethod),
                 C.byref(C.c_double(1e-10)), C.c_char_p(pc.encode()), _i(500), C.byref(ier),
                 C.c_size_t(len(vname)), C.c_size_t(len(cmethod)), C.c_size_t(4))
    it_f = gpu.iter_tot(2)
    assert ier.value == 0
    assert it_f > 0
    gpu.grid_init(im, jm, km)
    var2, am2, bm2 = np.zeros(b.size), a.copy(), b.copy()
    gpu.reset_iter_tot()
    ier2 = gpu.leq_bicgs("U_g", 2, var2, am2, bm2, "RSRS", 1e-10, pc, 500)
    assert ier2 == ier.value
    assert gpu.iter_tot(2) == it_f
    np.testing.assert_array_equal(var, var2)
    np.testing.assert_array_equal(am, am2)
    np.testing.assert_array_equal(bm, bm2)
