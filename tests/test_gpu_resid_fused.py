"""GPU: CALC_RESID + UNDER_RELAX_S + LEQ_BICGS in the solver's set-up pass
(mfix_leq_resid_relax_bicgs_dev) against the three entries run in turn.

MFIX runs CALC_RESID_S, UNDER_RELAX_S and SOLVE_LIN_EQ one after the other
(model/solve_scalar_eq.f:129-152); the fused entry forms the residual row and
applies the relaxation in the pass that scales the rows.  The relaxed
coefficients, the scaling and so the solve are the same arithmetic, so VAR,
the scaled A_M/B_M, IER and the iteration count must be bit-identical to
mfix_under_relax_s_dev followed by the solve, and the residual's fluid-cell
count, largest residual and its cell must equal mfix_calc_resid_*_dev's; NUM
and DEN are the same terms summed in another order (relative 1e-13).  The
separate entries are themselves checked against the oracle (test_gpu_resid,
test_gpu_phi, test_gpu_solver).  Cases: every residual kind, relaxation on
and off, FP64 and mixed, DIAG and LINE, 3-D, 2-D and cyclic grids, and a
grid of more than 2^18 cells (several finish blocks, the mixed restart
path).
"""
import numpy as np
import pytest

from mfix_amd import leq
from mfix_amd.leq import DeviceArray
from tests.cases import make_system
from tests.resid_cases import flags_for

pytestmark = pytest.mark.gpu

# (name, sys kind, imax, jmax, kmax, no_k, cyclic, resid kind, ur, pc, mixed)
CASES = [
    ("s_relax_diag", 0, 24, 20, 16, False, (0, 0, 0), "S", 0.8, "DIAG", False),
    ("s_relax_line_mixed", 1, 20, 34, 12, False, (0, 0, 0), "S", 0.5, "LINE", True),
    ("c_norelax_line", 1, 16, 18, 14, False, (1, 0, 0), "C", 0.0, "LINE", False),
    ("pp_diag", 0, 22, 18, 20, False, (0, 0, 0), "PP", 0.0, "DIAG", False),
    ("s_2d", 0, 40, 36, 1, True, (0, 0, 0), "S", 0.7, "DIAG", False),
    ("c_relax_cyc_mixed", 0, 18, 16, 14, False, (0, 0, 1), "C", 0.9, "DIAG", True),
    ("none_relax", 2, 20, 20, 20, False, (0, 0, 0), None, 0.6, "NONE", False),
    ("s_66cubed_mixed", 0, 66, 66, 66, False, (0, 0, 0), "S", 0.8, "DIAG", True),
    ("c_66cubed", 0, 66, 66, 66, False, (0, 0, 0), "C", 0.8, "DIAG", False),
]


def _inputs(case):
    name, sk, im, jm, km, no_k, cyc, rk, ur, pc, mixed = case
    a, b = make_system(sk, im, jm, km, no_k, seed=211)
    n3 = b.size
    rng = np.random.default_rng(17)
    x = rng.uniform(-1, 1, n3)
    flag = flags_for(im, jm, km, no_k, 99)
    ep_s = rng.uniform(0, 2e-4, n3) if rk == "S" else None
    return a, b, x, flag, ep_s


def _solve_kw(pc, mixed):
    return dict(pc=pc, cmethod="RSRS", tol=1e-9, itmax=400, mixed=mixed, seed=5)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fused_resid_relax_solve_equals_separate(gpu, case):
    name, sk, im, jm, km, no_k, cyc, rk, ur, pc, mixed = case
    print(f"fused residual + relaxation + solve: {name}", flush=True)
    gpu.grid_init(im, jm, km, no_k=no_k, cyclic=tuple(bool(c) for c in cyc))
    a, b, x, flag, ep_s = _inputs(case)
    dflag = DeviceArray.from_host(np.ascontiguousarray(flag, dtype=np.int32))
    deps = DeviceArray.from_host(ep_s) if ep_s is not None else None
    rkw = dict(tol=0.05, ep_s=deps, dil_ep_s=1.0e-4, norm=0.5, debug=True)

    # the reference's order: residual, relaxation, solve
    da, db, dx = DeviceArray.from_host(a), DeviceArray.from_host(b), DeviceArray.from_host(x)
    want_r = None
    if rk is not None:
        want_r = gpu.calc_resid_dev(rk, b=db, flag=dflag, var=dx, a=da, **rkw)
    if ur != 0.0:
        gpu.under_relax_s_dev(dx, dflag, ur, da, db)
    want = gpu.bicgs_dev(dx, da, db, **_solve_kw(pc, mixed))
    want_x, want_a, want_b = dx.download(), da.download(), db.download()

    # fused
    fa, fb, fx = DeviceArray.from_host(a), DeviceArray.from_host(b), DeviceArray.from_host(x)
    got_r, got = gpu.resid_relax_bicgs_dev(rk, fx, fa, fb, dflag, ur, resid_tol=rkw["tol"], ep_s=deps,
                                           dil_ep_s=rkw["dil_ep_s"], norm=rkw["norm"], debug=True,
                                           **_solve_kw(pc, mixed))
    assert got.ier == want.ier
    for k in ("iterations", "ier", "exit_reason", "rnorm0", "rnorm", "true_rnorm", "refinements"):
        assert got.stats[k] == want.stats[k], (k, got.stats[k], want.stats[k])
    np.testing.assert_array_equal(fx.download(), want_x)
    np.testing.assert_array_equal(fa.download(), want_a)
    np.testing.assert_array_equal(fb.download(), want_b)
    if rk is None:
        assert got_r is None
    else:
        assert got_r["ncells"] == want_r["ncells"] and got_r["ijk_resid"] == want_r["ijk_resid"], (got_r, want_r)
        for k in ("num", "den", "resid", "max_resid"):
            np.testing.assert_allclose(got_r[k], want_r[k], rtol=1e-13, atol=0.0, err_msg=k)
        assert got_r["ncells"] > 0


def test_fused_entry_rejects_bad_arguments(gpu):
    gpu.grid_init(6, 5, 4)
    a, b = make_system(0, 6, 5, 4, False, seed=3)
    da, db, dx = DeviceArray.from_host(a), DeviceArray.from_host(b), DeviceArray.from_host(np.zeros(b.size))
    dflag = DeviceArray.from_host(np.ones(b.size, dtype=np.int32))
    with pytest.raises(leq.SolverAbort):
        gpu.resid_relax_bicgs_dev("S", dx, da, db, dflag, 0.8, pc="XXXX")
    L = leq._lib_()
    rc = L.mfix_leq_resid_relax_bicgs_dev(7, dflag.ptr, 0.0, None, 0.0, 1.0, 1, None, 0.0, b"RSRS", 1e-9, b"DIAG",
                                          10, 0, dx.ptr, da.ptr, db.ptr, None, None, None)
    assert rc == -104   # MFIX_LEQ_EINVAL: unknown residual kind
