"""CPU oracle of the LEQ_BICGS path (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package: it checks the HIP solver and is never part
of the product path.  ``leq_oracle.c`` is a plain-C restatement of the
reference routines (each function cites the file:line it follows);
``ref_harness.f90`` drives the reference's own compiled routines
(``make ref``, only where the MFIX tree is present) to write golden vectors.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libleq_oracle.so"
REF_BIN = HERE / "_ref" / "leq_ref_harness"
REF_TREE = Path("/root/reference")

_dp = C.POINTER(C.c_double)


class OGrid(C.Structure):
    _fields_ = [
        ("imax", C.c_int), ("jmax", C.c_int), ("kmax", C.c_int), ("no_k", C.c_int),
        ("cyclic_x", C.c_int), ("cyclic_y", C.c_int), ("cyclic_z", C.c_int),
        ("is_serial", C.c_int), ("minimize_dotproducts", C.c_int), ("icheck_bicgs", C.c_int),
    ]


class OResult(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("iter_tot_increment", C.c_int), ("ier", C.c_int),
        ("exit_reason", C.c_int), ("rnorm0", C.c_double), ("rnorm", C.c_double),
        ("msolve_fatal", C.c_int),
    ]


class OResid(C.Structure):
    _fields_ = [
        ("num", C.c_double), ("den", C.c_double), ("resid", C.c_double), ("max_resid", C.c_double),
        ("ijk_resid", C.c_int), ("ncells", C.c_int),
    ]


RESID_KINDS = {"C": 0, "S": 1, "PP": 2}

# field order of the pressure-correction assembly (og_pp_fields, ref_harness op 10)
PP_FIELDS = ("rop_ge", "rop_gn", "rop_gt", "ayz", "axz", "axy", "vol", "rop_g", "rop_go", "u_g", "v_g", "w_g",
             "sum_r_g", "d_e", "d_n", "d_t")


class OPPFields(C.Structure):
    _fields_ = [(f, _dp) for f in PP_FIELDS] + [("flag", C.POINTER(C.c_int)), ("odt", C.c_double),
                                                ("ijk_p_g", C.c_int)]


# scalar transport (og_phi_fields, ref_harness op 11): DIMENSION_3 fields,
# then the 1-D geometry arrays along I, J and K (Fortran index 1.. at [0..])
PHI_FIELDS = ("dif", "u", "v", "w", "flux_e", "flux_n", "flux_t", "ayz", "axz", "axy")
PHI_GEO_I = ("odx_e", "ox", "fx_e", "fx_e_bar")
PHI_GEO_J = ("ody_n", "fy_n", "fy_n_bar")
PHI_GEO_K = ("odz_t", "fz_t", "fz_t_bar")
PHI_SOURCES = ("s_p", "s_c", "ep", "phi")


class OPhiFields(C.Structure):
    _fields_ = [(f, _dp) for f in PHI_FIELDS + PHI_GEO_I + PHI_GEO_J + PHI_GEO_K] + [
        ("flag", C.POINTER(C.c_int))]


def build_oracle(force: bool = False) -> Path:
    if force or not LIB.exists() or LIB.stat().st_mtime < (HERE / "leq_oracle.c").stat().st_mtime:
        subprocess.run(["make", "-s", "-C", str(HERE), "oracle"], check=True)
    return LIB


def build_ref() -> Path | None:
    """Compile the reference's LEQ_BICGS closure (only where the tree exists)."""
    if not (REF_TREE / "model" / "leq_bicgs.f").exists():
        return REF_BIN if REF_BIN.exists() else None
    subprocess.run(["make", "-s", "-C", str(HERE), "ref", f"REF={REF_TREE}"], check=True)
    return REF_BIN


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        build_oracle()
        l = C.CDLL(str(LIB))
        l.oracle_matvec.argtypes = [C.POINTER(OGrid), _dp, _dp, _dp]
        l.oracle_dgtsv.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        l.oracle_dgtsv.restype = C.c_int
        l.oracle_msolve_line.argtypes = [C.POINTER(OGrid), _dp, _dp, _dp, C.c_char_p]
        l.oracle_msolve_line.restype = C.c_int
        l.oracle_msolve_diag.argtypes = [C.POINTER(OGrid), _dp, _dp, _dp]
        l.oracle_msolve_none.argtypes = [C.POINTER(OGrid), _dp, _dp]
        l.oracle_dot.argtypes = [C.POINTER(OGrid), _dp, _dp]
        l.oracle_dot.restype = C.c_double
        l.oracle_dot2.argtypes = [C.POINTER(OGrid), _dp, _dp, _dp, _dp, _dp]
        l.oracle_unit_scale.argtypes = [C.POINTER(OGrid), _dp, _dp]
        l.oracle_unit_scale.restype = C.c_int
        l.oracle_leq_bicgs.argtypes = [C.POINTER(OGrid), C.c_char_p, C.c_char_p, C.c_double, C.c_int,
                                       _dp, _dp, _dp, _dp, C.POINTER(OResult), _dp]
        l.oracle_leq_bicgs.restype = C.c_int
        l.oracle_calc_resid.argtypes = [C.POINTER(OGrid), C.c_int, _dp, _dp, _dp, C.POINTER(C.c_int), C.c_double,
                                        _dp, C.c_double, C.c_double, C.c_int, C.POINTER(OResid)]
        l.oracle_assemble_pp_g.argtypes = [C.POINTER(OGrid), C.POINTER(OPPFields), _dp, _dp, _dp]
        l.oracle_assemble_pp_g.restype = C.c_int
        l.oracle_conv_dif_phi0.argtypes = [C.POINTER(OGrid), C.POINTER(OPhiFields), _dp, _dp]
        l.oracle_conv_dif_phi0.restype = None
        l.oracle_source_phi.argtypes = [C.POINTER(OGrid), _dp, _dp, _dp, _dp, C.POINTER(C.c_int), C.c_double,
                                        _dp, _dp]
        l.oracle_source_phi.restype = None
        l.oracle_under_relax_s.argtypes = [C.POINTER(OGrid), _dp, C.POINTER(C.c_int), C.c_double, _dp, _dp]
        l.oracle_under_relax_s.restype = None
        _lib = l
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def grid(imax, jmax, kmax, no_k=False, cyclic=(False, False, False), is_serial=True,
         minimize_dotproducts=False, icheck_bicgs=1) -> OGrid:
    return OGrid(imax, jmax, kmax, int(no_k), int(cyclic[0]), int(cyclic[1]), int(cyclic[2]),
                 int(is_serial), int(minimize_dotproducts), int(icheck_bicgs))


def n3_of(g: OGrid) -> int:
    return (g.imax + 2) * (g.jmax + 2) * (1 if g.no_k else g.kmax + 2)


def matvec(g, a, x):
    y = np.zeros(n3_of(g))
    lib().oracle_matvec(C.byref(g), _p(a), _p(x), _p(y))
    return y


def dgtsv(dl, d, du, b):
    dl, d, du, b = (np.array(v, dtype=np.float64) for v in (dl, d, du, b))
    info = lib().oracle_dgtsv(len(d), _p(dl), _p(d), _p(du), _p(b))
    return b, info


def msolve(g, pc, a, r, cmethod="RSRS"):
    z = np.zeros(n3_of(g))
    if pc == "LINE":
        info = lib().oracle_msolve_line(C.byref(g), _p(r), _p(a), _p(z), cmethod.encode())
        return z, info
    if pc == "DIAG":
        lib().oracle_msolve_diag(C.byref(g), _p(r), _p(a), _p(z))
    else:
        lib().oracle_msolve_none(C.byref(g), _p(r), _p(z))
    return z, 0


def dot(g, x, y):
    return lib().oracle_dot(C.byref(g), _p(x), _p(y))


def dot2(g, r1, r2, r3, r4):
    out = np.zeros(2)
    lib().oracle_dot2(C.byref(g), _p(r1), _p(r2), _p(r3), _p(r4), _p(out))
    return out


def unit_scale(g, a, b):
    a = a.copy(); b = b.copy()
    ier = lib().oracle_unit_scale(C.byref(g), _p(a), _p(b))
    return a, b, ier


def leq_bicgs(g, pc, cmethod, tol, itmax, var, a, b, noise, trace=False):
    """Oracle LEQ_BICGS; works on copies, returns (x, a_scaled, b_scaled, result dict[, trace])."""
    var = var.copy(); a = a.copy(); b = b.copy()
    res = OResult()
    tr = np.full(itmax + 1, np.nan) if trace else None
    rc = lib().oracle_leq_bicgs(C.byref(g), pc.encode(), cmethod.encode(), tol, itmax, _p(var), _p(a),
                                _p(b), _p(np.ascontiguousarray(noise, dtype=np.float64)), C.byref(res),
                                _p(tr) if trace else None)
    if rc != 0:
        raise ValueError(f"oracle: preconditioner option {pc!r} not found")
    out = {f: getattr(res, f) for f, _ in res._fields_}
    if trace:
        return var, a, b, out, tr
    return var, a, b, out


def calc_resid(g, kind, var, a, b, flag, tol=0.0, ep_s=None, dil_ep_s=1.0e-4, norm=1.0, debug=True) -> dict:
    """Oracle CALC_RESID_C / _S / _PP (kind "C", "S", "PP"); flag is MFIX's
    FLAG array (int32, DIMENSION_3), ep_s the phase's EP_s or None (M = 0)."""
    out = OResid()
    fl = np.ascontiguousarray(flag, dtype=np.int32)
    eps = None if ep_s is None else np.ascontiguousarray(ep_s, dtype=np.float64)
    lib().oracle_calc_resid(C.byref(g), RESID_KINDS[kind], _p(np.ascontiguousarray(var, dtype=np.float64)),
                            _p(np.ascontiguousarray(a, dtype=np.float64)),
                            _p(np.ascontiguousarray(b, dtype=np.float64)),
                            fl.ctypes.data_as(C.POINTER(C.c_int)), tol, _p(eps) if eps is not None else None,
                            dil_ep_s, norm, int(debug), C.byref(out))
    return {f: getattr(out, f) for f, _ in out._fields_}


def assemble_pp_g(g, fields: dict, flag, odt, ijk_p_g=0):
    """Oracle INIT_AB_M + CONV_PP_G + SOURCE_PP_G for the gas phase; returns
    (A_M, B_M, B_MMAX, number of fluid cells left with A0 = 0 and B != 0)."""
    n3 = n3_of(g)
    arrs = {k: np.ascontiguousarray(fields[k], dtype=np.float64) for k in PP_FIELDS}
    fl = np.ascontiguousarray(flag, dtype=np.int32)
    F = OPPFields(*[_p(arrs[k]) for k in PP_FIELDS], fl.ctypes.data_as(C.POINTER(C.c_int)), odt, int(ijk_p_g))
    a = np.zeros(7 * n3)
    b = np.zeros(n3)
    bm = np.zeros(n3)
    bad = lib().oracle_assemble_pp_g(C.byref(g), C.byref(F), _p(a), _p(b), _p(bm))
    return a, b, bm, bad


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def conv_dif_phi(g, fields: dict, flag):
    """Oracle INIT_AB_M + CONV_DIF_PHI (first-order upwinding, DISC = 0);
    fields holds PHI_FIELDS and the 1-D geometry arrays.  Returns (A_M, B_M)."""
    n3 = n3_of(g)
    keep = {k: _f64(fields[k]) for k in PHI_FIELDS + PHI_GEO_I + PHI_GEO_J + PHI_GEO_K}
    fl = np.ascontiguousarray(flag, dtype=np.int32)
    F = OPhiFields(*[_p(keep[k]) for k in PHI_FIELDS + PHI_GEO_I + PHI_GEO_J + PHI_GEO_K], _iptr(fl))
    a = np.zeros(7 * n3)
    b = np.zeros(n3)
    lib().oracle_conv_dif_phi0(C.byref(g), C.byref(F), _p(a), _p(b))
    return a, b


def source_phi(g, a, b, s_p, s_c, ep, phi, flag, dil_ep_s=1.0e-4):
    """Oracle SOURCE_PHI on copies of (A_M, B_M); returns the new pair."""
    a = _f64(a).copy(); b = _f64(b).copy()
    fl = np.ascontiguousarray(flag, dtype=np.int32)
    lib().oracle_source_phi(C.byref(g), _p(_f64(s_p)), _p(_f64(s_c)), _p(_f64(ep)), _p(_f64(phi)), _iptr(fl),
                            dil_ep_s, _p(a), _p(b))
    return a, b


def under_relax_s(g, a, b, var, flag, ur):
    """Oracle UNDER_RELAX_S on copies of (A_M, B_M); returns the new pair."""
    a = _f64(a).copy(); b = _f64(b).copy()
    fl = np.ascontiguousarray(flag, dtype=np.int32)
    lib().oracle_under_relax_s(C.byref(g), _p(_f64(var)), _iptr(fl), ur, _p(a), _p(b))
    return a, b


# ---- driver of the compiled reference (ref_harness.f90) --------------------
OPS = {"bicgs": 1, "matvec": 2, "msolve": 3, "msolve1": 4, "dot": 5, "dot2": 6,
       "resid_c": 7, "resid_s": 8, "resid_pp": 9, "assemble_pp": 10, "phi": 11,
       "extract": 12, "write_ab_m": 13}


def run_reference(op, g: OGrid, a, b, x, pc="DIAG", cmethod="RSRS", tol=1e-10, itmax=500, seed=12345,
                  workdir: Path | None = None, resid: dict | None = None, pp: dict | None = None,
                  phi: dict | None = None):
    """Run one reference routine through oracle/_ref/leq_ref_harness.  The
    residual ops take resid = {flag, debug, m, tol, norm, ep_s}, the
    pressure-correction assembly op pp = {flag, odt, ijk_p_g, fields}, the
    scalar transport op phi = {flag, stage, ur, fields} (fields: PHI_FIELDS,
    PHI_SOURCES and the 1-D geometry arrays; stage 1 CONV_DIF_PHI, 2
    SOURCE_PHI, 3 UNDER_RELAX_S on the a, b given).  The dump writers
    (extract: matrixExtract, write_ab_m: WRITE_AB_M) return the text files
    they wrote under "files"."""
    if not REF_BIN.exists():
        raise FileNotFoundError(REF_BIN)
    n3 = n3_of(g)
    workdir = Path(workdir or os.environ.get("TMPDIR", "/tmp"))
    fin = workdir / f"refin_{os.getpid()}.bin"
    fout = workdir / f"refout_{os.getpid()}.bin"
    hdr = np.array([OPS[op], g.imax, g.jmax, g.kmax, g.no_k, g.cyclic_x, g.cyclic_y, g.cyclic_z,
                    g.is_serial, g.minimize_dotproducts, g.icheck_bicgs, itmax, seed], dtype=np.int32)
    with open(fin, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.float64(tol).tobytes())
        f.write(pc.encode().ljust(4)[:4])
        f.write(cmethod.encode().ljust(4)[:4])
        f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(b, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(x, dtype=np.float64).tobytes())
        if OPS[op] == 10:
            f.write(np.array([int(pp["ijk_p_g"])], dtype=np.int32).tobytes())
            f.write(np.float64(pp["odt"]).tobytes())
            f.write(np.ascontiguousarray(pp["flag"], dtype=np.int32).tobytes())
            f.write(np.stack([np.ascontiguousarray(pp["fields"][k], dtype=np.float64) for k in PP_FIELDS],
                             axis=1).T.copy().tobytes())
        elif OPS[op] == 11:
            fd = phi["fields"]
            f.write(np.array([int(phi["stage"])], dtype=np.int32).tobytes())
            f.write(np.float64(phi.get("ur", 1.0)).tobytes())
            f.write(np.ascontiguousarray(phi["flag"], dtype=np.int32).tobytes())
            for keys in (PHI_FIELDS + PHI_SOURCES, PHI_GEO_I, PHI_GEO_J, PHI_GEO_K):
                f.write(np.stack([_f64(fd[k]) for k in keys]).tobytes())
        elif 7 <= OPS[op] <= 9:
            f.write(np.array([int(resid["debug"]), int(resid["m"])], dtype=np.int32).tobytes())
            f.write(np.array([resid.get("tol", 0.0), resid.get("norm", 1.0)], dtype=np.float64).tobytes())
            f.write(np.ascontiguousarray(resid["flag"], dtype=np.int32).tobytes())
            eps = resid.get("ep_s")
            f.write(np.ascontiguousarray(np.zeros(n3) if eps is None else eps, dtype=np.float64).tobytes())
    import tempfile
    files = {}
    try:
        with tempfile.TemporaryDirectory(dir=workdir) as cwd:
            subprocess.run([str(REF_BIN), str(fin), str(fout)], check=True, capture_output=True, cwd=cwd,
                           env={**os.environ, "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS", "1")})
            for q in Path(cwd).iterdir():
                files[q.name] = q.read_text()
        raw = fout.read_bytes()
    finally:
        for p in (fin, fout):
            try:
                p.unlink()
            except FileNotFoundError:
                pass
    ier, iter_tot = np.frombuffer(raw[:8], dtype=np.int32)
    off = 8
    dres = np.frombuffer(raw[off:off + 16], dtype=np.float64).copy(); off += 16
    y = np.frombuffer(raw[off:off + 8 * n3], dtype=np.float64).copy(); off += 8 * n3
    a_out = np.frombuffer(raw[off:off + 56 * n3], dtype=np.float64).copy(); off += 56 * n3
    b_out = np.frombuffer(raw[off:off + 8 * n3], dtype=np.float64).copy(); off += 8 * n3
    noise = np.frombuffer(raw[off:off + 8 * n3], dtype=np.float64).copy()
    return {"ier": int(ier), "iter_tot": int(iter_tot), "dot": dres, "x": y, "a": a_out, "b": b_out,
            "noise": noise, "files": files}
