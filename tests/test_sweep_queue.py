"""Deferred SOLVE_LIN_EQ calls (mfix_solve_lin_eq_deferred) without a GPU:
the checks SOLVE_LIN_EQ and LEQ_BICGS make at the call (METHOD 2 only,
model/solve_lin_eq.f:139-190; PC LINE/DIAG/NONE, model/leq_bicgs.f:83-87)
happen when a call is queued, nothing touches the device until the flush,
and the queue counts and discards calls."""
from types import SimpleNamespace

import pytest

from mfix_amd import leq
from mfix_amd._lib import LeqError


def _fake(n):
    # queued pointers are not dereferenced before the flush
    return SimpleNamespace(ptr=0x1000 + 8 * n)


@pytest.fixture
def grid(native_lib):
    leq.grid_init(8, 7, 6)
    leq.sweep_discard()
    yield
    leq.sweep_discard()


def test_queue_counts_and_discards(grid):
    for q in range(5):
        leq.solve_lin_eq_deferred("U_g", 3, _fake(3 * q), _fake(3 * q + 1), _fake(3 * q + 2), 5)
    assert leq.sweep_pending() == 5
    leq.sweep_discard()
    assert leq.sweep_pending() == 0


def test_rejections_happen_at_the_call(grid):
    with pytest.raises(leq.SolverAbort):
        leq.solve_lin_eq_deferred("U_g", 3, _fake(0), _fake(1), _fake(2), 5, pc="ILU0")
    with pytest.raises(LeqError) as e:
        leq.solve_lin_eq_deferred("U_g", 3, _fake(0), _fake(1), _fake(2), 5, method=1)
    assert e.value.code == -108
    with pytest.raises(LeqError):
        leq.solve_lin_eq_deferred("U_g", 0, _fake(0), _fake(1), _fake(2), 5)     # VNO outside 1..10
    with pytest.raises(LeqError):
        leq.solve_lin_eq_deferred("U_g", 3, SimpleNamespace(ptr=None), _fake(1), _fake(2), 5)
    assert leq.sweep_pending() == 0


def test_queue_capacity(grid):
    for q in range(1024):
        leq.solve_lin_eq_deferred("X", 9, _fake(q), _fake(q), _fake(q), 1)
    with pytest.raises(LeqError) as e:
        leq.solve_lin_eq_deferred("X", 9, _fake(0), _fake(0), _fake(0), 1)
    assert e.value.code == -107
    assert leq.sweep_pending() == 1024
