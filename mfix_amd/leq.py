"""Host-side mirror of the reference's LEQ interface, backed by the HIP library.

``leq_bicgs`` and ``solve_lin_eq`` take the reference's arguments in the
reference's order (model/leq_bicgs.f:23-24, model/solve_lin_eq.f:17-18) and
act on numpy arrays in the reference's serial storage; like the reference
they overwrite VAR with the solution and A_M/B_M with the row-scaled system
and return IER (0, -1 or -2).  An unknown PC raises, where the reference calls
mfix_exit (model/leq_bicgs.f:83-87).

``DeviceArray`` and the ``*_dev`` functions keep A_M/B_M/VAR resident in HBM
for repeated solves; every call goes through libmfix_leq.so, and nothing in
this module computes a result on the host.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import BatchItem, Grid, Opts, Resid, Stats, check


class SolverAbort(RuntimeError):
    """Raised where the reference aborts through mfix_exit."""


def _lib_():
    return _lib.load()


def init(device: int = 0) -> None:
    check(_lib_().mfix_leq_init(device), "mfix_leq_init")


def finalize() -> None:
    check(_lib_().mfix_leq_finalize(), "mfix_leq_finalize")


def grid_init(imax: int, jmax: int, kmax: int, *, no_k: bool = False, cyclic=(False, False, False),
              is_serial: bool = True, minimize_dotproducts: bool = False, icheck_bicgs: int = 1) -> Grid:
    """Replaces the GRIDMAP_INIT/SET_INCREMENTS module state the solver reads."""
    g = Grid(imax, jmax, kmax, int(no_k), int(cyclic[0]), int(cyclic[1]), int(cyclic[2]),
             int(is_serial), int(minimize_dotproducts), int(icheck_bicgs))
    check(_lib_().mfix_leq_grid_init(C.byref(g)), "mfix_leq_grid_init")
    return g


def dimension_3() -> int:
    return check(_lib_().mfix_leq_dimension_3(), "mfix_leq_dimension_3")


def iter_tot(vno: int) -> int:
    return check(_lib_().mfix_leq_iter_tot(vno), "mfix_leq_iter_tot")


def reset_iter_tot() -> None:
    _lib_().mfix_leq_reset_iter_tot()


def _dp(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.c_contiguous or not a.flags.writeable:
        raise TypeError("arrays must be writable, C-contiguous float64")
    return a.ctypes.data_as(_lib.c_double_p)


def set_noise(noise: np.ndarray | None, seed: int = 0) -> None:
    """Inject the U[0,1) array used for Rtilde by ``leq_bicgs`` (parity runs)."""
    global _noise_keepalive
    _noise_keepalive = None if noise is None else np.ascontiguousarray(noise, dtype=np.float64)
    ptr = None if noise is None else _noise_keepalive.ctypes.data
    check(_lib_().mfix_leq_set_noise(ptr, C.c_uint64(seed)), "mfix_leq_set_noise")


_noise_keepalive = None


def leq_bicgs(vname: str, vno: int, var: np.ndarray, a_m: np.ndarray, b_m: np.ndarray, cmethod: str,
              tol: float, pc: str, itmax: int) -> int:
    """LEQ_BICGS (model/leq_bicgs.f:23-90) on host arrays; returns IER."""
    ier = C.c_int(0)
    n3 = dimension_3()
    if var.size != n3 or a_m.size != 7 * n3 or b_m.size != n3:
        raise ValueError(f"array sizes do not match DIMENSION_3={n3}")
    vb = vname.encode()
    cb = cmethod.encode()
    rc = _lib_().mfix_leq_bicgs(vb, len(vb), vno, _dp(var), _dp(a_m), _dp(b_m), cb, len(cb), float(tol),
                                pc.encode().ljust(4)[:4], int(itmax), C.byref(ier))
    if rc == -101:
        raise SolverAbort("preconditioner option not found - check mfix.dat and readme")
    if rc == -106:
        raise SolverAbort("DGTSV failed in a JK/IJ line sweep")
    check(rc, "mfix_leq_bicgs")
    return ier.value


def last_stats() -> dict:
    st = Stats()
    check(_lib_().mfix_leq_last_stats(C.byref(st)), "mfix_leq_last_stats")
    return st.as_dict()


def solve_lin_eq(vname: str, vno: int, var: np.ndarray, a_m: np.ndarray, b_m: np.ndarray, m: int,
                 itmax: int, method: int, sweep: str, tol1: float, pc: str) -> int:
    """SOLVE_LIN_EQ (model/solve_lin_eq.f:17-195): a_m is (M+1, 7, N3), b_m (M+1, N3)."""
    ier = C.c_int(0)
    vb = vname.encode()
    rc = _lib_().mfix_solve_lin_eq(vb, len(vb), vno, _dp(var), _dp(a_m), _dp(b_m), int(m), int(itmax),
                                   int(method), sweep.encode().ljust(4)[:4], float(tol1),
                                   pc.encode().ljust(4)[:4], C.byref(ier))
    if rc == -101:
        raise SolverAbort("preconditioner option not found - check mfix.dat and readme")
    if rc == -108:
        raise SolverAbort(f"LEQ_METHOD = {method} is not provided by this library")
    check(rc, "mfix_solve_lin_eq")
    return ier.value


# ---- device-resident arrays ---------------------------------------------------

class DeviceArray:
    """A block of HBM owned by libmfix_leq (float64 or float32 elements)."""

    def __init__(self, n: int, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        check(_lib_().mfix_leq_dev_alloc(self.nbytes, C.byref(p)), "mfix_leq_dev_alloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(_lib_().mfix_leq_copy_to_dev(self.ptr, a.ctypes.data, self.nbytes), "copy_to_dev")

    def download(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        check(_lib_().mfix_leq_copy_from_dev(out.ctypes.data, self.ptr, self.nbytes), "copy_from_dev")
        return out

    def copy_from(self, other: "DeviceArray") -> None:
        assert other.nbytes == self.nbytes
        check(_lib_().mfix_leq_copy_dev(self.ptr, other.ptr, self.nbytes), "copy_dev")

    def free(self) -> None:
        if self.ptr:
            _lib_().mfix_leq_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.free()
        except Exception:
            pass


@dataclass
class SolveResult:
    ier: int
    stats: dict


def _opts(noise: DeviceArray | None, seed: int) -> Opts:
    o = Opts()
    o.write_back_scaled = 1
    o.noise = noise.ptr if noise is not None else None
    o.noise_seed = seed
    return o


def bicgs_dev(var: DeviceArray, a_m: DeviceArray, b_m: DeviceArray, *, pc: str = "DIAG",
              cmethod: str = "RSRS", tol: float = 1e-10, itmax: int = 500, noise: DeviceArray | None = None,
              seed: int = 0, mixed: bool = False) -> SolveResult:
    """FP64 (or mixed FP32/FP64) LEQ_BICGS on resident device arrays."""
    st = Stats()
    o = _opts(noise, seed)
    fn = _lib_().mfix_leq_bicgs_mixed_dev if mixed else _lib_().mfix_leq_bicgs_dev
    rc = fn(cmethod.encode().ljust(4)[:4], float(tol), pc.encode().ljust(4)[:4], int(itmax), var.ptr, a_m.ptr,
            b_m.ptr, C.byref(o), C.byref(st), None)
    if rc == -101:
        raise SolverAbort("preconditioner option not found - check mfix.dat and readme")
    check(rc, "bicgs_dev")
    return SolveResult(st.ier, st.as_dict())


def matvec_dev(a_m: DeviceArray, x: DeviceArray, y: DeviceArray) -> None:
    """y = A x.  FP64: LEQ_MATVEC on A_M in the reference layout
    (model/leqsol_mod.f:103-265).  FP32: the mixed inner solve's row on the
    difference form (plane 3 holds the row excess A0 + sum of the couplings)."""
    if a_m.dtype == np.float32:
        check(_lib_().mfix_leq_matvec32_dev(a_m.ptr, x.ptr, y.ptr, None), "matvec32")
    else:
        check(_lib_().mfix_leq_matvec_dev(a_m.ptr, x.ptr, y.ptr, None), "matvec")


def msolve_dev(pc: str, cmethod: str, a_m: DeviceArray, r: DeviceArray, z: DeviceArray) -> int:
    rc = _lib_().mfix_leq_msolve_dev(pc.encode().ljust(4)[:4], cmethod.encode().ljust(4)[:4], a_m.ptr, r.ptr,
                                     z.ptr, None)
    if rc == -106:
        raise SolverAbort("DGTSV failed in a JK/IJ line sweep")
    return check(rc, "msolve")


def msolve32_dev(cmethod: str, a_m: DeviceArray, r: DeviceArray, z: DeviceArray) -> int:
    """LINE preconditioner of the mixed inner solve on an 8-plane FP32
    matrix (couplings at planes 0-2 and 4-6, A0 at plane 7)."""
    assert a_m.dtype == np.float32 and r.dtype == np.float32 and z.dtype == np.float32
    rc = _lib_().mfix_leq_msolve32_dev(cmethod.encode().ljust(4)[:4], a_m.ptr, r.ptr, z.ptr, None)
    if rc == -106:
        raise SolverAbort("DGTSV failed in a JK/IJ line sweep")
    return check(rc, "msolve32")


def dot_dev(x: DeviceArray, y: DeviceArray) -> float:
    out = C.c_double()
    check(_lib_().mfix_leq_dot_dev(x.ptr, y.ptr, C.byref(out), None), "dot")
    return out.value


def scale_dev(a_m: DeviceArray, b_m: DeviceArray) -> int:
    ier = C.c_int()
    check(_lib_().mfix_leq_scale_dev(a_m.ptr, b_m.ptr, C.byref(ier), None), "scale")
    return ier.value


def prof_enable(which: int, on: bool = True) -> None:
    check(_lib_().mfix_leq_prof_enable(which, int(on)), "prof_enable")


def prof_read(which: int) -> dict:
    n = C.c_int64()
    ms = C.c_double()
    by = C.c_double()
    check(_lib_().mfix_leq_prof_read(which, C.byref(n), C.byref(ms), C.byref(by)), "prof_read")
    ni = C.c_int64()
    mi = C.c_double()
    check(_lib_().mfix_leq_prof_idle(which, C.byref(ni), C.byref(mi)), "prof_idle")
    return {"launches": n.value, "total_ms": ms.value, "alg_bytes": by.value, "idle_launches": ni.value,
            "idle_ms": mi.value}


def set_graphs(on: bool) -> None:
    """Run iterations as captured HIP-graph replays (default) or direct launches."""
    check(_lib_().mfix_leq_set_graphs(int(on)), "set_graphs")


def set_line_block(imax: int, jmax: int, kmax: int, lines_per_block: int, mixed: bool = False) -> None:
    """Lines per block (64, 32, 16; 0 = size default) of the red-black line
    kernel for LINE solves of this grid size and precision."""
    check(_lib_().mfix_leq_set_line_block(imax, jmax, kmax, int(mixed), lines_per_block), "set_line_block")


def set_mixed_delta(delta: float) -> None:
    """Reliable-update trigger of mixed solves that set none: the FP32
    recursive residual fell by `delta` (0 = the size default)."""
    check(_lib_().mfix_leq_set_mixed_delta(float(delta)), "set_mixed_delta")


def set_unit_diag(on: bool) -> None:
    """Mixed DIAG solves negate instead of reading 1/A0 when every 1/A0 is -1
    (default on); results are bit-identical either way."""
    check(_lib_().mfix_leq_set_unit_diag(int(on)), "set_unit_diag")


# ---- batched solves -------------------------------------------------------------

@dataclass
class BatchSystem:
    """One system of a batch: its grid, solver settings and resident arrays."""
    imax: int
    jmax: int
    kmax: int
    var: DeviceArray
    a_m: DeviceArray
    b_m: DeviceArray
    pc: str = "DIAG"
    cmethod: str = "RSRS"
    tol: float = 1e-10
    itmax: int = 500
    mixed: bool = False
    seed: int = 0


def _pad4(s: str) -> bytes:
    return s.encode().ljust(4)[:4]


def batch_solve(systems: list[BatchSystem]) -> list[SolveResult]:
    """Solve independent systems together: one launch per phase covers all of
    them (mfix_leq_batch_solve).  Each result equals the system's single
    solve bit for bit."""
    n = len(systems)
    items = (BatchItem * max(n, 1))()
    for q, s in enumerate(systems):
        it = items[q]
        it.imax, it.jmax, it.kmax = s.imax, s.jmax, s.kmax
        it.pc = _pad4(s.pc)
        it.cmethod = _pad4(s.cmethod)
        it.tol = float(s.tol)
        it.itmax = int(s.itmax)
        it.mixed = int(bool(s.mixed))
        it.d_var, it.d_a_m, it.d_b_m = s.var.ptr, s.a_m.ptr, s.b_m.ptr
        it.noise_seed = int(s.seed)
    stats = (Stats * max(n, 1))()
    rc = _lib_().mfix_leq_batch_solve(n, items, stats, None)
    if rc == -101:
        raise SolverAbort("preconditioner option not found - check mfix.dat and readme")
    check(rc, "mfix_leq_batch_solve")
    return [SolveResult(stats[q].ier, stats[q].as_dict()) for q in range(n)]


# ---- residual norms: CALC_RESID_C / _S / _PP (model/calc_resid.f) ----------

def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def calc_resid(kind: str, *, b, flag, var=None, a=None, tol: float = 0.0, ep_s=None, dil_ep_s: float = 1.0e-4,
               norm: float = 1.0, debug: bool = True) -> dict:
    """CALC_RESID_C ("C"), CALC_RESID_S ("S") or CALC_RESID_PP ("PP") of
    host arrays (model/calc_resid.f:16, :195, :391): var = VAR, a = A_M of
    the phase, b = B_M of the phase, flag = MFIX's FLAG (int32), ep_s = the
    solids phase's EP_s (S with M > 0) or None.  Returns NUM, DEN, RESID,
    MAX_RESID, IJK_RESID (1-based) and the fluid-cell count."""
    out = Resid()
    fl = np.ascontiguousarray(flag, dtype=np.int32)
    L = _lib_()
    if kind == "PP":
        check(L.mfix_calc_resid_pp(_f64(b).ctypes.data, fl.ctypes.data, norm, int(debug), C.byref(out)),
              "mfix_calc_resid_pp")
    elif kind == "C":
        check(L.mfix_calc_resid_c(_f64(var).ctypes.data, _f64(a).ctypes.data, _f64(b).ctypes.data, fl.ctypes.data,
                                  int(debug), C.byref(out)), "mfix_calc_resid_c")
    elif kind == "S":
        eps = None if ep_s is None else _f64(ep_s)
        check(L.mfix_calc_resid_s(_f64(var).ctypes.data, _f64(a).ctypes.data, _f64(b).ctypes.data, fl.ctypes.data,
                                  tol, None if eps is None else eps.ctypes.data, dil_ep_s, int(debug), C.byref(out)),
              "mfix_calc_resid_s")
    else:
        raise ValueError(f"unknown residual kind {kind!r}")
    return out.as_dict()


def calc_resid_dev(kind: str, *, b: "DeviceArray", flag: "DeviceArray", var: "DeviceArray | None" = None,
                   a: "DeviceArray | None" = None, tol: float = 0.0, ep_s: "DeviceArray | None" = None,
                   dil_ep_s: float = 1.0e-4, norm: float = 1.0, debug: bool = True) -> dict:
    """calc_resid on device-resident arrays (flag: an int32 DeviceArray)."""
    out = Resid()
    L = _lib_()
    if kind == "PP":
        check(L.mfix_calc_resid_pp_dev(b.ptr, flag.ptr, norm, int(debug), C.byref(out), None), "mfix_calc_resid_pp_dev")
    elif kind == "C":
        check(L.mfix_calc_resid_c_dev(var.ptr, a.ptr, b.ptr, flag.ptr, int(debug), C.byref(out), None),
              "mfix_calc_resid_c_dev")
    elif kind == "S":
        check(L.mfix_calc_resid_s_dev(var.ptr, a.ptr, b.ptr, flag.ptr, tol, None if ep_s is None else ep_s.ptr,
                                      dil_ep_s, int(debug), C.byref(out), None), "mfix_calc_resid_s_dev")
    else:
        raise ValueError(f"unknown residual kind {kind!r}")
    return out.as_dict()


_RESID_KINDS = {"C": 0, "S": 1, "PP": 2}


def resid_relax_bicgs_dev(kind: str | None, var: "DeviceArray", a_m: "DeviceArray", b_m: "DeviceArray",
                          flag: "DeviceArray", ur: float = 0.0, *, resid_tol: float = 0.0,
                          ep_s: "DeviceArray | None" = None, dil_ep_s: float = 1.0e-4, norm: float = 1.0,
                          debug: bool = True, pc: str = "DIAG", cmethod: str = "RSRS", tol: float = 1e-10,
                          itmax: int = 500, mixed: bool = False, noise: "DeviceArray | None" = None,
                          seed: int = 0) -> tuple[dict | None, "SolveResult"]:
    """CALC_RESID (kind "C", "S", "PP" or None), UNDER_RELAX_S (ur, 0 = none)
    and LEQ_BICGS of one equation, the residual and the relaxation done by
    the solver's set-up pass.  Returns (the residual's dict or None, the
    solve's result)."""
    out = Resid()
    st = Stats()
    o = _opts(noise, seed)
    k = -1 if kind is None else _RESID_KINDS[kind]
    rc = _lib_().mfix_leq_resid_relax_bicgs_dev(
        k, flag.ptr, float(resid_tol), None if ep_s is None else ep_s.ptr, float(dil_ep_s), float(norm), int(debug),
        C.byref(out), float(ur), cmethod.encode().ljust(4)[:4], float(tol), pc.encode().ljust(4)[:4], int(itmax),
        int(mixed), var.ptr, a_m.ptr, b_m.ptr, C.byref(o), C.byref(st), None)
    if rc == -101:
        raise SolverAbort("preconditioner option not found - check mfix.dat and readme")
    check(rc, "resid_relax_bicgs_dev")
    return (None if k < 0 else out.as_dict()), SolveResult(st.ier, st.as_dict())


# ---- the gas pressure-correction equation on the device ---------------------
# (INIT_AB_M + CONV_PP_G + SOURCE_PP_G, SOLVE_PP_G: model/conv_pp_g.f,
#  model/source_pp_g.f, model/solve_pp_g.f)

def _pp_struct(fields: dict, flag: "DeviceArray", odt: float, ijk_p_g: int) -> "_lib.PPFields":
    f = _lib.PPFields()
    for k in _lib.PP_FIELDS:
        setattr(f, k, fields[k].ptr)
    f.flag = flag.ptr
    f.odt = float(odt)
    f.ijk_p_g = int(ijk_p_g)
    return f


def assemble_pp_g_dev(fields: dict, flag: "DeviceArray", odt: float, ijk_p_g: int, a: "DeviceArray",
                      b: "DeviceArray", b_mmax: "DeviceArray") -> int:
    """Assemble A_M(:,:,0), B_M(:,0) and B_MMAX(:,0) of the gas pressure
    correction from resident fields (DeviceArrays keyed by the names of
    mfix_pp_g_fields); returns the count of fluid cells with A0 = 0, B != 0."""
    f = _pp_struct(fields, flag, odt, ijk_p_g)
    nbad = C.c_int()
    check(_lib_().mfix_assemble_pp_g_dev(C.byref(f), a.ptr, b.ptr, b_mmax.ptr, C.byref(nbad), None),
          "mfix_assemble_pp_g_dev")
    return nbad.value


def solve_pp_g_dev(fields: dict, flag: "DeviceArray", odt: float, ijk_p_g: int, normg: float, pp_g: "DeviceArray",
                   a: "DeviceArray", b: "DeviceArray", b_mmax: "DeviceArray", *, itmax: int = 500, method: int = 2,
                   sweep: str = "RSRS", tol: float = 1e-10, pc: str = "DIAG", mixed: bool = False) -> dict:
    """SOLVE_PP_G for the gas phase: returns RESG, IER and the solver stats."""
    f = _pp_struct(fields, flag, odt, ijk_p_g)
    resg = C.c_double()
    ier = C.c_int()
    st = Stats()
    check(_lib_().mfix_solve_pp_g_dev(C.byref(f), normg, pp_g.ptr, a.ptr, b.ptr, b_mmax.ptr, itmax, method,
                                      sweep.encode(), tol, pc.encode(), int(mixed), C.byref(resg), C.byref(ier),
                                      C.byref(st), None), "mfix_solve_pp_g_dev")
    return {"resg": resg.value, "ier": ier.value, "stats": st.as_dict()}


# ---- scalar transport equations on the device (first-order upwinding) ------
# (INIT_AB_M + CONV_DIF_PHI, SOURCE_PHI, UNDER_RELAX_S: model/conv_dif_phi.f,
#  model/source_phi.f, model/under_relax.f)

def conv_dif_phi_dev(fields: dict, flag: "DeviceArray", a: "DeviceArray", b: "DeviceArray", *,
                     sync: bool = True) -> None:
    """INIT_AB_M + CONV_DIF_PHI (DISC = 0) into a (7 planes) and b from resident
    fields (DeviceArrays keyed by the names of mfix_phi_fields)."""
    f = _lib.PhiFields()
    for k in _lib.PHI_FIELDS + _lib.PHI_GEO:
        setattr(f, k, fields[k].ptr)
    f.flag = flag.ptr
    L = _lib_()
    check(L.mfix_conv_dif_phi_dev(C.byref(f), a.ptr, b.ptr, None), "mfix_conv_dif_phi_dev")
    if sync:
        check(L.mfix_leq_sync(), "mfix_leq_sync")


def source_phi_dev(s_p: "DeviceArray", s_c: "DeviceArray", ep: "DeviceArray", phi: "DeviceArray",
                   flag: "DeviceArray", a: "DeviceArray", b: "DeviceArray", *, dil_ep_s: float = 1.0e-4,
                   sync: bool = True) -> None:
    """SOURCE_PHI in place on (a, b)."""
    L = _lib_()
    check(L.mfix_source_phi_dev(s_p.ptr, s_c.ptr, ep.ptr, phi.ptr, flag.ptr, dil_ep_s, a.ptr, b.ptr, None),
          "mfix_source_phi_dev")
    if sync:
        check(L.mfix_leq_sync(), "mfix_leq_sync")


def under_relax_s_dev(var: "DeviceArray", flag: "DeviceArray", ur: float, a: "DeviceArray", b: "DeviceArray", *,
                      sync: bool = True) -> None:
    """UNDER_RELAX_S in place on (a, b)."""
    L = _lib_()
    check(L.mfix_under_relax_s_dev(var.ptr, flag.ptr, ur, a.ptr, b.ptr, None), "mfix_under_relax_s_dev")
    if sync:
        check(L.mfix_leq_sync(), "mfix_leq_sync")


# ---- SIMPLE-sweep batching (deferred SOLVE_LIN_EQ, model/solve_lin_eq.f) ----

class Deferred:
    """A queued SOLVE_LIN_EQ call; ier and stats are filled at the flush."""

    def __init__(self, vname: str, vno: int, var: "DeviceArray", a_m: "DeviceArray", b_m: "DeviceArray"):
        self.vname, self.vno = vname, vno
        self.arrays = (var, a_m, b_m)   # kept alive until the flush
        self._ier = C.c_int(0)
        self.stats: dict | None = None

    @property
    def ier(self) -> int:
        return self._ier.value


_queued: list[Deferred] = []


def solve_lin_eq_deferred(vname: str, vno: int, var: "DeviceArray", a_m: "DeviceArray", b_m: "DeviceArray",
                          itmax: int, sweep: str = "RSRS", tol: float = 1e-4, pc: str = "LINE", *,
                          method: int = 2, mixed: bool = False) -> Deferred:
    """Queue SOLVE_LIN_EQ(VNAME, VNO, VAR, A_M(:,:,M), B_M(:,M), ...) on
    resident arrays for the next sweep_flush (argument checks happen now)."""
    d = Deferred(vname, vno, var, a_m, b_m)
    rc = _lib_().mfix_solve_lin_eq_deferred(vname.encode(), len(vname), vno, var.ptr, a_m.ptr, b_m.ptr, itmax,
                                            method, _pad4(sweep), tol, _pad4(pc), int(mixed), C.byref(d._ier))
    if rc == -101:
        raise SolverAbort("preconditioner option not found - check mfix.dat and readme")
    check(rc, "mfix_solve_lin_eq_deferred")
    _queued.append(d)
    return d


def sweep_pending() -> int:
    return _lib_().mfix_leq_sweep_pending()


def sweep_discard() -> None:
    _lib_().mfix_leq_sweep_discard()
    _queued.clear()


def sweep_flush() -> list[Deferred]:
    """Solve the queued calls together; returns them with ier and stats set."""
    n = sweep_pending()
    st = (Stats * max(n, 1))()
    done = list(_queued)
    _queued.clear()
    rc = _lib_().mfix_leq_sweep_flush(st, None)
    if rc in (0, -106):   # with a DGTSV failure every IER is still written
        for q, d in enumerate(done[:n]):
            d.stats = st[q].as_dict()
    check(rc, "mfix_leq_sweep_flush")
    return done
