"""SIMPLE-sweep batching on the GPU: SOLVE_LIN_EQ calls queued with
mfix_solve_lin_eq_deferred and solved together at mfix_leq_sweep_flush give,
call by call, the single device solve's result bit for bit (the batched
systems run the same kernels with their own state), the same IER and the
same ITER_TOT(VNO) accounting (model/leqsol_mod.f:17), on 3-D, 2-D and
cyclic grids and with FP64 and mixed-precision calls in one flush."""
import numpy as np
import pytest

from mfix_amd import leq, systems
from mfix_amd.leq import DeviceArray
from tests.cases import iterations_ok, oracle_spread

pytestmark = pytest.mark.gpu

# a SOLVE_VEL_STAR-like stage: U, V, W of the gas (VNO 3, 4, 5), one solids
# phase's U (VNO 3) and a scalar (VNO 9) and a pressure-like system
STAGE = [("U_g", 3, 1, "LINE", "RSRS", 5, 1e-4, False), ("V_g", 4, 1, "LINE", "RSRS", 5, 1e-4, False),
         ("W_g", 5, 1, "LINE", "RSRS", 5, 1e-4, False), ("U_s", 3, 1, "LINE", "ISIS", 5, 1e-4, True),
         ("Scalar 1", 9, 2, "DIAG", "RSRS", 50, 1e-8, False), ("P_like", 1, 0, "DIAG", "RSRS", 200, 1e-10, True)]


def _grid_system(kind, im, jm, km, no_k, seed):
    if no_k:
        from tests.cases import make_system
        return make_system(kind, im, jm, 1, no_k=True, seed=seed)
    return systems.gen_system(kind, im, jm, km, seed, nspheres=6)


@pytest.mark.parametrize("dims,no_k,cyc", [((20, 18, 16), False, (0, 0, 0)), ((40, 30, 1), True, (0, 0, 0)),
                                           ((18, 16, 14), False, (1, 0, 1))])
def test_flush_equals_sequential_solves(gpu, dims, no_k, cyc):
    im, jm, km = dims
    print(f"sweep flush vs single solves {dims} no_k={no_k} cyc={cyc}", flush=True)
    gpu.grid_init(im, jm, km, no_k=no_k, cyclic=tuple(bool(c) for c in cyc))
    sysdata = [_grid_system(kind, im, jm, km, no_k, 700 + q) for q, (_, _, kind, *_rest) in enumerate(STAGE)]
    # sequential: one device solve per call
    gpu.reset_iter_tot()
    seq = []
    for (vname, vno, kind, pc, sweep, itmax, tol, mixed), (a, b) in zip(STAGE, sysdata):
        da, db, dx = DeviceArray.from_host(a), DeviceArray.from_host(b), DeviceArray.from_host(np.zeros(b.size))
        r = gpu.bicgs_dev(dx, da, db, pc=pc, cmethod=sweep, tol=tol, itmax=itmax, mixed=mixed)
        seq.append((dx.download(), r))
    # the device entry has no VNO: ITER_TOT(VNO) is what LEQ_BICGS would add
    tot_seq = {v: 0 for v in (1, 3, 4, 5, 9)}
    for (vname, vno, *_rest), (_, r) in zip(STAGE, seq):
        tot_seq[vno] += r.stats["iter_tot_increment"]
    assert any(tot_seq.values())
    # deferred: queued, then one flush
    gpu.reset_iter_tot()
    held, calls = [], []
    for (vname, vno, kind, pc, sweep, itmax, tol, mixed), (a, b) in zip(STAGE, sysdata):
        da, db, dx = DeviceArray.from_host(a), DeviceArray.from_host(b), DeviceArray.from_host(np.zeros(b.size))
        held.append(dx)
        calls.append(gpu.solve_lin_eq_deferred(vname, vno, dx, da, db, itmax, sweep, tol, pc, mixed=mixed))
    assert gpu.sweep_pending() == len(STAGE)
    done = gpu.sweep_flush()
    assert gpu.sweep_pending() == 0
    for (x_seq, r_seq), d, dx in zip(seq, done, held):
        assert d.ier == r_seq.ier
        assert d.stats["iterations"] == r_seq.stats["iterations"]
        np.testing.assert_array_equal(dx.download(), x_seq)
    assert {v: gpu.iter_tot(v) for v in (1, 3, 4, 5, 9)} == tot_seq


def test_flush_of_nothing_and_reuse(gpu):
    gpu.grid_init(12, 10, 8)
    assert gpu.sweep_flush() == []
    a, b = systems.gen_system(1, 12, 10, 8, 3, nspheres=6)
    for rep in range(2):   # the queue empties at a flush and takes new calls
        da, db, dx = DeviceArray.from_host(a), DeviceArray.from_host(b), DeviceArray.from_host(np.zeros(b.size))
        gpu.solve_lin_eq_deferred("U_g", 3, dx, da, db, 20, "RSRS", 1e-8, "LINE")
        (d,) = gpu.sweep_flush()
        assert d.ier == 0 and d.stats["iterations"] > 0


def test_each_call_keeps_its_grid(gpu):
    """A flush solves every queued call on the grid that was current at the
    call, so calls on different grids can share one flush."""
    grids = [(12, 10, 8), (9, 14, 11)]
    singles, held = [], []
    for q, (im, jm, km) in enumerate(grids):
        a, b = systems.gen_system(1, im, jm, km, 40 + q, nspheres=6)
        gpu.grid_init(im, jm, km)
        da, db, dx = DeviceArray.from_host(a), DeviceArray.from_host(b), DeviceArray.from_host(np.zeros(b.size))
        singles.append((dx, gpu.bicgs_dev(dx, da, db, pc="LINE", cmethod="RSRS", tol=1e-8, itmax=50)))
        da, db, dx = DeviceArray.from_host(a), DeviceArray.from_host(b), DeviceArray.from_host(np.zeros(b.size))
        held.append(dx)
        gpu.solve_lin_eq_deferred("U_g", 3, dx, da, db, 50, "RSRS", 1e-8, "LINE")
    done = gpu.sweep_flush()
    for (x_single, r), d, dx in zip(singles, done, held):
        assert d.ier == r.ier and d.stats["iterations"] == r.stats["iterations"]
        np.testing.assert_array_equal(dx.download(), x_single.download())


@pytest.mark.parametrize("dims,no_k,cyc", [((20, 18, 16), False, (0, 0, 0)), ((40, 30, 1), True, (0, 0, 0)),
                                           ((18, 16, 14), False, (1, 0, 1))])
def test_flush_against_oracle(gpu, oracle_mod, dims, no_k, cyc):
    """Each call of a flushed stage against the oracle's LEQ_BICGS on the
    same inputs and R~ noise (a deferred call draws the library's default
    stream, seed 0): FP64 calls give the same IER and iterations within 10 %
    (or within the oracle's own rounding spread, tests.cases.oracle_spread);
    a converged FP64 call's solution is within 1e-8, and one stopped at
    ITMAX differs from the oracle's iterate no more than the oracle's
    iterates differ among themselves under rounding-level perturbations
    (x4); mixed calls that converge meet TOL on the FP64 true residual."""
    im, jm, km = dims
    print(f"sweep flush vs oracle {dims} no_k={no_k} cyc={cyc}", flush=True)
    gpu.grid_init(im, jm, km, no_k=no_k, cyclic=tuple(bool(c) for c in cyc))
    og = oracle_mod.grid(im, jm, km, no_k=no_k, cyclic=cyc)
    sysdata = [_grid_system(kind, im, jm, km, no_k, 800 + q) for q, (_, _, kind, *_rest) in enumerate(STAGE)]
    held, calls = [], []
    for (vname, vno, kind, pc, sweep, itmax, tol, mixed), (a, b) in zip(STAGE, sysdata):
        da, db, dx = DeviceArray.from_host(a), DeviceArray.from_host(b), DeviceArray.from_host(np.zeros(b.size))
        held.append((dx, da, db))
        calls.append(gpu.solve_lin_eq_deferred(vname, vno, dx, da, db, itmax, sweep, tol, pc, mixed=mixed))
    done = gpu.sweep_flush()
    for (vname, vno, kind, pc, sweep, itmax, tol, mixed), (a, b), d, (dx, da, db) in zip(STAGE, sysdata, done, held):
        n3 = b.size
        noise = systems.gen_uniform(n3, 0, 100)
        xo, ao, bo, res = oracle_mod.leq_bicgs(og, pc, sweep, tol, itmax, np.zeros(n3), a, b, noise)
        x = dx.download()
        np.testing.assert_array_equal(da.download(), ao)
        scale = max(np.max(np.abs(xo)), 1e-300)
        if not mixed:
            assert d.ier == res["ier"], (vname, d.stats, res)
            spread = None
            if not iterations_ok(d.stats["iterations"], res["iterations"], []):
                spread = oracle_spread(oracle_mod, og, pc, sweep, tol, itmax, np.zeros(n3), a, b, noise)
                assert iterations_ok(d.stats["iterations"], res["iterations"], spread), vname
            if res["exit_reason"] == 0:
                assert np.max(np.abs(x - xo)) <= 1e-8 * scale, vname
            else:
                spread = spread or oracle_spread(oracle_mod, og, pc, sweep, tol, itmax, np.zeros(n3), a, b, noise)
                own = max(np.max(np.abs(xs - xo)) for xs, _ in spread)
                assert np.max(np.abs(x - xo)) <= 4.0 * own + 1e-8 * scale, (vname, own / scale)
        elif d.ier == 0:
            r = bo - oracle_mod.matvec(og, ao, x)
            assert np.sqrt(np.dot(r, r)) <= tol * res["rnorm0"] * 1.0000001, vname
            if res["ier"] == 0:
                assert np.max(np.abs(x - xo)) <= max(1e-8, 50.0 * tol) * scale, vname
