"""mfix_leq_bicgs_dev_noalloc (the entry point that allocates nothing)
against mfix_leq_bicgs_dev, through the C ABI, on the GPU.

Both run the same kernels in the same order on a table of one, so the
results must be the same bits: VAR, the row-scaled A_M/B_M, IER and the
iteration count.  Two calls in a row, the second on a smaller system, check
that nothing of the first call's contents in the shared fixed array leaks
into the second; a system whose workspace does not fit is refused with
MFIX_LEQ_ETOOBIG before anything runs.
"""
import ctypes as C

import numpy as np
import pytest

from mfix_amd import _lib, leq, systems
from mfix_amd.leq import DeviceArray

pytestmark = pytest.mark.gpu

ETOOBIG = -109


def _noalloc(lib, dx, da, db, pc, cm, tol, itmax, noise=None):
    lib.mfix_leq_bicgs_dev_noalloc.restype = C.c_int
    lib.mfix_leq_bicgs_dev_noalloc.argtypes = [C.c_char_p, C.c_double, C.c_char_p, C.c_int, C.c_void_p, C.c_void_p,
                                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    st = _lib.Stats()
    o = _lib.Opts()
    o.write_back_scaled = 1
    o.noise = noise.ptr if noise is not None else None
    rc = lib.mfix_leq_bicgs_dev_noalloc(cm.encode(), tol, pc.encode(), itmax, dx.ptr, da.ptr, db.ptr, C.byref(o),
                                        C.byref(st), None)
    return rc, st.as_dict()


def _run_both(gpu, kind, n, pc, cm, seed, tol=1e-10, itmax=2000):
    lib = _lib.load()
    gpu.grid_init(n, n + 3, n - 2)
    a, b = systems.gen_system(kind, n, n + 3, n - 2, seed, nspheres=8)
    noise = DeviceArray.from_host(np.random.default_rng(seed).random(b.size))
    da, db, dx = DeviceArray.from_host(a), DeviceArray.from_host(b), DeviceArray.from_host(np.zeros(b.size))
    ref = gpu.bicgs_dev(dx, da, db, pc=pc, cmethod=cm, tol=tol, itmax=itmax, noise=noise)
    x_ref, a_ref, b_ref = dx.download(), da.download(), db.download()
    da2, db2, dx2 = DeviceArray.from_host(a), DeviceArray.from_host(b), DeviceArray.from_host(np.zeros(b.size))
    rc, st = _noalloc(lib, dx2, da2, db2, pc, cm, tol, itmax, noise)
    assert rc == 0, (rc, _lib.load().mfix_leq_error_string(rc))
    assert st["ier"] == ref.ier
    assert st["iterations"] == ref.stats["iterations"] and st["iter_tot_increment"] == ref.stats["iter_tot_increment"]
    np.testing.assert_array_equal(dx2.download(), x_ref)
    np.testing.assert_array_equal(da2.download(), a_ref)
    np.testing.assert_array_equal(db2.download(), b_ref)
    return st


@pytest.mark.parametrize("pc,cm,kind", [("DIAG", "RSRS", 0), ("LINE", "RSRS", 1), ("NONE", "RSRS", 0),
                                        ("LINE", "ISIS", 1)])
def test_noalloc_matches_allocating_entry_twice(gpu, pc, cm, kind):
    print(f"no-allocation entry vs mfix_leq_bicgs_dev, {pc} {cm}", flush=True)
    first = _run_both(gpu, kind, 40, pc, cm, 11)   # 42 x 45 x 40 = 75,600 cells
    second = _run_both(gpu, kind, 18, pc, cm, 12)  # smaller: 20 x 23 x 18
    assert first["ier"] == 0 and second["ier"] == 0


def test_noalloc_refuses_a_system_that_does_not_fit(gpu):
    lib = _lib.load()
    n = 140  # 142 x 145 x 140 = 2.9 M cells: more than 64 MiB of FP64 workspace
    gpu.grid_init(n, n + 3, n - 2)
    n3 = (n + 2) * (n + 5) * n
    dx, db, da = DeviceArray(n3), DeviceArray(n3), DeviceArray(7 * n3)
    rc, _ = _noalloc(lib, dx, da, db, "DIAG", "RSRS", 1e-10, 10)
    assert rc == ETOOBIG
    lib.mfix_leq_error_string.restype = C.c_char_p
    assert b"no-allocation" in lib.mfix_leq_error_string(rc)
    # the refusal left the entry usable
    _run_both(gpu, 0, 16, "DIAG", "RSRS", 13)
