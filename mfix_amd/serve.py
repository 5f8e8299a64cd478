"""Serving front end: continuous batching over libmfix_leq's request loop.

``Server`` wraps mfix_leq_server_*: requests (host A_M/B_M/VAR of any grid
size) queue in arrival order; every ``step`` admits what fits in the free
slots and the HBM budget, advances all admitted systems together by one
chunk of iterations (one launch per phase for all of them) and retires the
finished ones, writing the solution into the request's VAR.  The host
arrays of a request are kept referenced here until it completes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import Request, Stats, check, load


@dataclass
class Done:
    ier: int
    stats: dict
    var: np.ndarray


class Server:
    def __init__(self, hbm_budget_bytes: int, max_slots: int = 64):
        self._lib = load()
        h = C.c_void_p()
        check(self._lib.mfix_leq_server_create(int(hbm_budget_bytes), int(max_slots), C.byref(h)),
              "mfix_leq_server_create")
        self._h = h.value
        self._held: dict[int, tuple] = {}

    def submit(self, a_m: np.ndarray, b_m: np.ndarray, var: np.ndarray, imax: int, jmax: int, kmax: int, *,
               pc: str = "DIAG", cmethod: str = "RSRS", tol: float = 1e-10, itmax: int = 500, mixed: bool = False,
               seed: int = 0) -> int:
        """Queue a solve; A_M and B_M are copied at admission, VAR receives
        the solution when the request retires."""
        for arr in (a_m, b_m, var):
            if arr.dtype != np.float64 or not arr.flags.c_contiguous:
                raise ValueError("A_M, B_M and VAR must be C-contiguous float64 arrays")
        n3 = (imax + 2) * (jmax + 2) * (kmax + 2)
        if a_m.size != 7 * n3 or b_m.size != n3 or var.size != n3:
            raise ValueError("array sizes do not match the grid")
        r = Request()
        r.imax, r.jmax, r.kmax = imax, jmax, kmax
        r.pc = pc.encode().ljust(4)[:4]
        r.cmethod = cmethod.encode().ljust(4)[:4]
        r.tol = float(tol)
        r.itmax = int(itmax)
        r.mixed = int(bool(mixed))
        r.a_m, r.b_m, r.var = a_m.ctypes.data, b_m.ctypes.data, var.ctypes.data
        r.noise_seed = int(seed)
        rid = self._lib.mfix_leq_server_submit(self._h, C.byref(r))
        check(int(rid) if rid < 0 else 0, "mfix_leq_server_submit")
        self._held[int(rid)] = (a_m, b_m, var)
        return int(rid)

    def step(self) -> None:
        check(self._lib.mfix_leq_server_step(self._h), "mfix_leq_server_step")

    def poll(self, rid: int) -> Done | None:
        st = Stats()
        rc = self._lib.mfix_leq_server_poll(self._h, int(rid), C.byref(st))
        check(rc if rc < 0 else 0, "mfix_leq_server_poll")
        if rc == 0:
            return None
        _, _, var = self._held.pop(int(rid))
        return Done(st.ier, st.as_dict(), var)

    def pending(self) -> int:
        return check(self._lib.mfix_leq_server_pending(self._h), "mfix_leq_server_pending")

    def run_until_done(self, max_rounds: int = 1_000_000) -> None:
        for _ in range(max_rounds):
            if self.pending() == 0:
                return
            self.step()
        raise RuntimeError("server did not drain")

    def close(self) -> None:
        if self._h:
            self._lib.mfix_leq_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
