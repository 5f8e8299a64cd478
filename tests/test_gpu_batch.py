"""Batched and served solves against single solves (GPU).

A batched system runs the same kernels with its own state, partial sums and
block split as the single solve, so its result must be the single solve's
bit for bit; single solves are themselves checked against the oracle in
test_gpu_solver.py.  The same holds for a request served alone or in a mix
of requests admitted at different rounds, for graph replays against direct
launches, and for a budget so small that requests are admitted one by one.
"""
import numpy as np
import pytest

from mfix_amd import leq, systems
from mfix_amd.leq import BatchSystem, DeviceArray
from mfix_amd.serve import Server

pytestmark = pytest.mark.gpu

# (kind, imax, jmax, kmax, pc, cmethod, mixed)
MIX = [
    (0, 12, 10, 9, "DIAG", "RSRS", False),
    (1, 9, 14, 11, "LINE", "RSRS", False),
    (0, 20, 18, 16, "NONE", "RSRS", False),
    (1, 16, 16, 16, "LINE", "ISIS", False),   # one-by-one sweep path
    (2, 24, 20, 22, "DIAG", "RSRS", True),
    (1, 13, 11, 15, "LINE", "RSSS", True),
    (0, 7, 9, 8, "DIAG", "RSRS", False),
]


def _system(kind, im, jm, km, seed):
    a, b = systems.gen_system(kind, im, jm, km, seed, nspheres=6)
    return a, b


def _single(gpu, kind, im, jm, km, pc, cm, mixed, seed, tol=1e-10, itmax=400):
    a, b = _system(kind, im, jm, km, seed)
    gpu.grid_init(im, jm, km)
    da, db = DeviceArray.from_host(a), DeviceArray.from_host(b)
    dx = DeviceArray.from_host(np.zeros(b.size))
    res = gpu.bicgs_dev(dx, da, db, pc=pc, cmethod=cm, tol=tol, itmax=itmax, seed=seed, mixed=mixed)
    return dx.download(), res


def _batch(gpu, cases, tol=1e-10, itmax=400):
    held, systems_ = [], []
    for q, (kind, im, jm, km, pc, cm, mixed) in enumerate(cases):
        a, b = _system(kind, im, jm, km, 500 + q)
        da, db = DeviceArray.from_host(a), DeviceArray.from_host(b)
        dx = DeviceArray.from_host(np.zeros(b.size))
        held.append((da, db, dx))
        systems_.append(BatchSystem(im, jm, km, dx, da, db, pc=pc, cmethod=cm, tol=tol, itmax=itmax, mixed=mixed,
                                    seed=500 + q))
    res = leq.batch_solve(systems_)
    return [h[2].download() for h in held], res


def test_batch_equals_single_solves(gpu):
    print("batch vs single, 7 systems", flush=True)
    xs, res = _batch(gpu, MIX)
    for q, (kind, im, jm, km, pc, cm, mixed) in enumerate(MIX):
        x1, r1 = _single(gpu, kind, im, jm, km, pc, cm, mixed, 500 + q)
        assert res[q].ier == r1.ier == 0, (q, res[q].stats, r1.stats)
        assert res[q].stats["iterations"] == r1.stats["iterations"], q
        np.testing.assert_array_equal(xs[q], x1, err_msg=f"system {q}")


def test_batch_graph_replays_match_direct_launches(gpu):
    """Captured chunks replayed many times, with eager work in between,
    give the direct-launch result on every call."""
    cases = MIX[:3] + MIX[4:5]
    leq.set_graphs(False)
    try:
        ref, rref = _batch(gpu, cases)
    finally:
        leq.set_graphs(True)
    for rep in range(3):
        print(f"graph replay round {rep}", flush=True)
        xs, res = _batch(gpu, cases)
        for q in range(len(cases)):
            assert res[q].stats["iterations"] == rref[q].stats["iterations"]
            np.testing.assert_array_equal(xs[q], ref[q], err_msg=f"rep {rep} system {q}")
        # eager work between replays: a direct single solve
        _single(gpu, 0, 8, 8, 8, "DIAG", "RSRS", False, 7)


def test_single_solve_graph_and_direct_agree(gpu):
    leq.set_graphs(False)
    try:
        x0, r0 = _single(gpu, 1, 15, 13, 14, "LINE", "RSRS", False, 11)
    finally:
        leq.set_graphs(True)
    for _ in range(2):
        x1, r1 = _single(gpu, 1, 15, 13, 14, "LINE", "RSRS", False, 11)
        np.testing.assert_array_equal(x1, x0)
        assert r1.stats["iterations"] == r0.stats["iterations"]


def test_batch_of_one_and_empty(gpu):
    assert leq.batch_solve([]) == []
    xs, res = _batch(gpu, MIX[:1])
    x1, _ = _single(gpu, *MIX[0][:4], *MIX[0][4:7], 500)
    np.testing.assert_array_equal(xs[0], x1)


def test_batch_rejects_bad_pc(gpu):
    a, b = _system(0, 4, 4, 4, 1)
    da, db, dx = DeviceArray.from_host(a), DeviceArray.from_host(b), DeviceArray.from_host(np.zeros(b.size))
    with pytest.raises(leq.SolverAbort):
        leq.batch_solve([BatchSystem(4, 4, 4, dx, da, db, pc="ILU0")])


def _serve(cases, budget=1 << 30, max_slots=16, stagger=0):
    """Submit `cases` (kind, side..., pc, cm, mixed, seed, tol); with stagger>0
    submit one request every `stagger` rounds so later ones join running
    ones.  Returns solutions and stats in submission order."""
    out = {}
    with Server(budget, max_slots) as srv:
        ids = []
        pending = list(cases)
        rounds = 0
        while pending or srv.pending():
            if pending and (stagger == 0 or rounds % stagger == 0):
                batch = pending if stagger == 0 else pending[:1]
                for (kind, im, jm, km, pc, cm, mixed, seed, tol) in batch:
                    a, b = _system(kind, im, jm, km, seed)
                    ids.append(srv.submit(a, b, np.zeros(b.size), im, jm, km, pc=pc, cmethod=cm, tol=tol,
                                          itmax=400, mixed=mixed, seed=seed))
                pending = pending[len(batch):]
            srv.step()
            rounds += 1
            assert rounds < 100000
        for q, rid in enumerate(ids):
            d = srv.poll(rid)
            assert d is not None
            out[q] = (d.var.copy(), d.ier, d.stats)
    return [out[q] for q in range(len(cases))]


SERVE = [(kind, im, jm, km, pc, cm, mixed, 900 + q, tol)
         for q, ((kind, im, jm, km, pc, cm, mixed), tol) in enumerate(zip(MIX, [1e-10, 1e-8] * 4))]


def test_served_result_same_alone_and_in_mix(gpu):
    print("serving: mix submitted at once", flush=True)
    mix = _serve(SERVE)
    print("serving: staggered arrivals", flush=True)
    stag = _serve(SERVE, stagger=3)
    for q, case in enumerate(SERVE):
        print(f"serving: request {q} alone", flush=True)
        alone = _serve([case])[0]
        assert mix[q][1] == alone[1] == 0
        np.testing.assert_array_equal(mix[q][0], alone[0], err_msg=f"request {q} mix vs alone")
        np.testing.assert_array_equal(stag[q][0], alone[0], err_msg=f"request {q} staggered vs alone")
        assert mix[q][2]["iterations"] == alone[2]["iterations"]


def test_served_equals_single_solve(gpu):
    res = _serve(SERVE[:3])
    for q, (kind, im, jm, km, pc, cm, mixed, seed, tol) in enumerate(SERVE[:3]):
        x1, r1 = _single(gpu, kind, im, jm, km, pc, cm, mixed, seed, tol=tol)
        np.testing.assert_array_equal(res[q][0], x1)


def test_small_budget_admits_one_at_a_time(gpu):
    """A 5 MB budget holds one of these requests at a time (a 32^3 FP64 DIAG
    request needs about 120 bytes per cell); every request still completes
    with the result it gets with room for all.  A request larger than the
    whole budget is refused at submission."""
    cases = [(0, 30, 30, 30, "DIAG", "RSRS", False, 41, 1e-10), (1, 28, 28, 28, "DIAG", "RSRS", False, 42, 1e-10),
             (0, 26, 26, 26, "NONE", "RSRS", False, 43, 1e-8)]
    res = _serve(cases, budget=5_000_000, max_slots=8)
    ref = _serve(cases)
    for q in range(len(cases)):
        np.testing.assert_array_equal(res[q][0], ref[q][0])
    with Server(1 << 20, 4) as srv:
        a, b = _system(0, 40, 40, 40, 1)
        with pytest.raises(Exception):
            srv.submit(a, b, np.zeros(b.size), 40, 40, 40)
