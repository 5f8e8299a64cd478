"""C ABI and host-side logic, without a GPU.

* libmfix_leq.so loads and exports every function include/mfix_leq.h
  declares; the Fortran shim exports the reference's LEQ_BICGS link name;
* host-only entries work without a device (grid descriptor, DIMENSION_3,
  ITER_TOT, error strings, PC validation before any device work, the
  synthetic-system generators and their determinism);
* the Python mirror raises where the reference aborts.
"""
import ctypes as C
import hashlib
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "mfix_leq.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(mfix_\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_path(native_lib):
    names = declared_functions()
    for must in ("mfix_leq_bicgs", "mfix_solve_lin_eq", "mfix_leq_grid_init", "mfix_leq_matvec_dev",
                 "mfix_leq_msolve_dev", "mfix_leq_dot_dev", "mfix_leq_bicgs_mixed_dev", "mfix_leq_batch_solve",
                 "mfix_leq_server_submit"):
        assert must in names


def test_library_exports_every_declared_symbol(native_lib):
    missing = [n for n in declared_functions() if not hasattr(native_lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_fortran_shim_exports_reference_link_name():
    from mfix_amd import FORTRAN_LIB_PATH, build
    if not FORTRAN_LIB_PATH.exists():
        if build.build_fortran() is None:
            pytest.skip("no flang in this image")
    lib = C.CDLL(str(FORTRAN_LIB_PATH))
    for sym in ("leq_bicgs_", "mfix_leq_grid_init_f_", "mfix_leq_iter_tot_f_"):
        assert hasattr(lib, sym)


def test_grid_descriptor_host_only(native_lib):
    from mfix_amd import leq
    leq.grid_init(32, 30, 28)
    assert leq.dimension_3() == 34 * 32 * 30
    leq.grid_init(10, 8, 99, no_k=True)
    assert leq.dimension_3() == 12 * 10
    g = leq.Grid(0, 4, 4, 0, 0, 0, 0, 1, 0, 1)
    assert native_lib.mfix_leq_grid_init(C.byref(g)) == -102
    huge = leq.Grid(2000, 2000, 2000, 0, 0, 0, 0, 1, 0, 1)
    assert native_lib.mfix_leq_grid_init(C.byref(huge)) == -104


def test_iter_tot_and_error_strings(native_lib):
    from mfix_amd import leq
    leq.reset_iter_tot()
    assert all(leq.iter_tot(v) == 0 for v in range(1, 11))
    assert native_lib.mfix_leq_iter_tot(0) == -104
    assert b"preconditioner" in native_lib.mfix_leq_error_string(-101)


def test_bad_pc_rejected_before_device_work(native_lib):
    """LEQ_BICGS aborts on an unknown PC (leq_bicgs.f:83-87) before computing."""
    from mfix_amd import leq
    leq.grid_init(4, 4, 4)
    n3 = leq.dimension_3()
    with pytest.raises(leq.SolverAbort):
        leq.leq_bicgs("X", 1, np.zeros(n3), np.ones(7 * n3), np.ones(n3), "RSRS", 1e-8, "ILU0", 10)
    with pytest.raises(leq.SolverAbort):
        leq.leq_bicgs("X", 1, np.zeros(n3), np.ones(7 * n3), np.ones(n3), "RSRS", 1e-8, "diag", 10)


def test_solve_lin_eq_rejects_other_methods(native_lib):
    from mfix_amd import leq
    leq.grid_init(4, 4, 4)
    n3 = leq.dimension_3()
    with pytest.raises(leq.SolverAbort):
        leq.solve_lin_eq("X", 1, np.zeros(n3), np.ones((1, 7, n3)), np.ones((1, n3)), 0, 10, 3, "RSRS", 1e-8,
                         "DIAG")


def test_generator_deterministic_and_well_formed(native_lib):
    from mfix_amd import systems
    for kind in (0, 1, 2):
        a1, b1 = systems.gen_system(kind, 9, 7, 6, 123, nspheres=5)
        a2, b2 = systems.gen_system(kind, 9, 7, 6, 123, nspheres=5)
        assert hashlib.sha256(a1.tobytes()).digest() == hashlib.sha256(a2.tobytes()).digest()
        assert np.array_equal(b1, b2)
        A = a1.reshape(7, 8, 11, 9)
        # ghost rows identity-like, interior rows: off-diagonals >= 0, A0 < 0
        assert np.all(A[3, 0] == -1.0) and np.all(A[:, 0][[0, 1, 2, 4, 5, 6]] == 0.0)
        inner = A[:, 1:-1, 1:-1, 1:-1]
        assert np.all(inner[[0, 1, 2, 4, 5, 6]] >= 0.0) and np.all(inner[3] < 0.0)
        # diagonally dominant
        assert np.all(-inner[3] >= inner[[0, 1, 2, 4, 5, 6]].sum(axis=0))


def test_symmetric_kinds_are_symmetric(native_lib):
    from mfix_amd import systems
    from tests.cases import to_csr
    for kind in (0, 2):
        a, _ = systems.gen_system(kind, 7, 6, 5, 5, nspheres=5)
        m = to_csr(a, 7, 6, 5).toarray()
        assert np.allclose(m, m.T, rtol=0, atol=0)


def test_momentum_flux_is_divergence_free(native_lib):
    """Upwind coefficients from a discretely divergence-free flux: for every
    interior cell sum(A_nb) - sum(outflow-upwind parts) balances, i.e. the
    row sum of the convective part is zero, so -A0 = sum(nb) + rho V/dt."""
    from mfix_amd import systems
    a, _ = systems.gen_system(1, 8, 7, 6, 9)
    A = a.reshape(7, 8, 10, 9)[:, 1:-1, 1:-1, 1:-1]
    transient = -(A[3] + A[[0, 1, 2, 4, 5, 6]].sum(axis=0))
    assert np.allclose(transient, transient.flat[0], rtol=1e-12)
    assert transient.flat[0] > 0.0


def test_config_systems_match_fixture_hashes(native_lib):
    """Bytes of the config-0 system are pinned by the golden fixture."""
    gold = np.load(ROOT / "tests" / "golden" / "leq_ref_golden.npz", allow_pickle=False)
    from mfix_amd import systems
    a, b = systems.config_system(0)
    assert hashlib.sha256(a.tobytes()).hexdigest() == str(gold["bicgs/config0/a_sha"])
