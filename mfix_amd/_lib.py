"""ctypes binding of libmfix_leq.so (the C ABI in include/mfix_leq.h).

The library is loaded from the package directory only; a missing library is
an error (there is no fallback implementation).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libmfix_leq.so"

c_int_p = C.POINTER(C.c_int)
c_double_p = C.POINTER(C.c_double)


class Grid(C.Structure):
    """mfix_leq_grid: replaces the compar/geometry/indices module state."""
    _fields_ = [
        ("imax", C.c_int), ("jmax", C.c_int), ("kmax", C.c_int), ("no_k", C.c_int),
        ("cyclic_x", C.c_int), ("cyclic_y", C.c_int), ("cyclic_z", C.c_int),
        ("is_serial", C.c_int), ("minimize_dotproducts", C.c_int), ("icheck_bicgs", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("iter_tot_increment", C.c_int), ("ier", C.c_int),
        ("exit_reason", C.c_int), ("rnorm0", C.c_double), ("rnorm", C.c_double),
        ("true_rnorm", C.c_double), ("refinements", C.c_int), ("inner_iterations", C.c_int),
        ("device_ms", C.c_double), ("restarts", C.c_int),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class Opts(C.Structure):
    _fields_ = [
        ("write_back_scaled", C.c_int), ("noise", C.c_void_p), ("noise_seed", C.c_uint64),
        ("inner_tol", C.c_double), ("max_refine", C.c_int),
    ]


class Resid(C.Structure):
    _fields_ = [
        ("num", C.c_double), ("den", C.c_double), ("resid", C.c_double), ("max_resid", C.c_double),
        ("ijk_resid", C.c_int), ("ncells", C.c_int),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


PP_FIELDS = ("rop_ge", "rop_gn", "rop_gt", "ayz", "axz", "axy", "vol", "rop_g", "rop_go", "u_g", "v_g", "w_g",
             "sum_r_g", "d_e", "d_n", "d_t")


class PPFields(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in PP_FIELDS] + [("flag", C.c_void_p), ("odt", C.c_double),
                                                      ("ijk_p_g", C.c_int)]


# mfix_phi_fields order (include/mfix_leq.h): DIMENSION_3 fields, then the
# 1-D geometry arrays along I, J and K
PHI_FIELDS = ("dif", "u", "v", "w", "flux_e", "flux_n", "flux_t", "ayz", "axz", "axy")
PHI_GEO = ("odx_e", "ox", "fx_e", "fx_e_bar", "ody_n", "fy_n", "fy_n_bar", "odz_t", "fz_t", "fz_t_bar")


class PhiFields(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in PHI_FIELDS + PHI_GEO] + [("flag", C.c_void_p)]


class BatchItem(C.Structure):
    _fields_ = [
        ("imax", C.c_int), ("jmax", C.c_int), ("kmax", C.c_int),
        ("pc", C.c_char * 4), ("cmethod", C.c_char * 4),
        ("tol", C.c_double), ("itmax", C.c_int), ("mixed", C.c_int),
        ("d_var", C.c_void_p), ("d_a_m", C.c_void_p), ("d_b_m", C.c_void_p),
        ("noise_seed", C.c_uint64),
    ]


class Request(C.Structure):
    _fields_ = [
        ("imax", C.c_int), ("jmax", C.c_int), ("kmax", C.c_int),
        ("pc", C.c_char * 4), ("cmethod", C.c_char * 4),
        ("tol", C.c_double), ("itmax", C.c_int), ("mixed", C.c_int),
        ("a_m", C.c_void_p), ("b_m", C.c_void_p), ("var", C.c_void_p),
        ("noise_seed", C.c_uint64),
    ]


_SIGS = {
    "mfix_leq_init": (C.c_int, [C.c_int]),
    "mfix_leq_finalize": (C.c_int, []),
    "mfix_leq_grid_init": (C.c_int, [C.POINTER(Grid)]),
    "mfix_leq_dimension_3": (C.c_int64, []),
    "mfix_leq_iter_tot": (C.c_int, [C.c_int]),
    "mfix_leq_reset_iter_tot": (None, []),
    "mfix_leq_error_string": (C.c_char_p, [C.c_int]),
    "mfix_leq_bicgs": (C.c_int, [C.c_char_p, C.c_int, C.c_int, c_double_p, c_double_p, c_double_p,
                                 C.c_char_p, C.c_int, C.c_double, C.c_char_p, C.c_int, c_int_p]),
    "mfix_solve_lin_eq": (C.c_int, [C.c_char_p, C.c_int, C.c_int, c_double_p, c_double_p, c_double_p,
                                    C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_double, C.c_char_p,
                                    c_int_p]),
    "mfix_leq_bicgs_dev": (C.c_int, [C.c_char_p, C.c_double, C.c_char_p, C.c_int, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.POINTER(Opts), C.POINTER(Stats),
                                     C.c_void_p]),
    "mfix_leq_bicgs_mixed_dev": (C.c_int, [C.c_char_p, C.c_double, C.c_char_p, C.c_int, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.POINTER(Opts),
                                           C.POINTER(Stats), C.c_void_p]),
    "mfix_leq_matvec_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mfix_leq_msolve_dev": (C.c_int, [C.c_char_p, C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "mfix_leq_dot_dev": (C.c_int, [C.c_void_p, C.c_void_p, c_double_p, C.c_void_p]),
    "mfix_leq_scale_dev": (C.c_int, [C.c_void_p, C.c_void_p, c_int_p, C.c_void_p]),
    "mfix_leq_matvec32_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mfix_leq_gen_system": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int,
                                      c_double_p, c_double_p]),
    "mfix_leq_gen_uniform": (C.c_int, [C.c_int64, C.c_uint64, C.c_uint64, c_double_p]),
    "mfix_leq_batch_solve": (C.c_int, [C.c_int, C.POINTER(BatchItem), C.POINTER(Stats),
                                       C.c_void_p]),
    "mfix_leq_autotune": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, c_int_p, c_int_p]),
    "mfix_leq_server_create": (C.c_int, [C.c_size_t, C.c_int, C.POINTER(C.c_void_p)]),
    "mfix_leq_server_submit": (C.c_int64, [C.c_void_p, C.POINTER(Request)]),
    "mfix_leq_server_step": (C.c_int, [C.c_void_p]),
    "mfix_leq_server_poll": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(Stats)]),
    "mfix_leq_server_pending": (C.c_int, [C.c_void_p]),
    "mfix_leq_server_destroy": (C.c_int, [C.c_void_p]),
    "mfix_leq_set_graphs": (C.c_int, [C.c_int]),
    "mfix_leq_set_unit_diag": (C.c_int, [C.c_int]),
    "mfix_leq_set_shape": (C.c_int, [C.c_int] * 6),
    "mfix_leq_set_line_block": (C.c_int, [C.c_int] * 5),
    "mfix_leq_set_mixed_delta": (C.c_int, [C.c_double]),
    "mfix_assemble_pp_g_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mfix_solve_pp_g_dev": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                      C.c_int, C.c_char_p, C.c_double, C.c_char_p, C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]),
    "mfix_solve_lin_eq_deferred": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_int, C.c_int, C.c_char_p, C.c_double, C.c_char_p, C.c_int,
                                             C.c_void_p]),
    "mfix_leq_sweep_flush": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mfix_leq_sweep_pending": (C.c_int, []),
    "mfix_leq_sweep_discard": (None, []),
    "mfix_conv_dif_phi_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mfix_source_phi_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                      C.c_void_p, C.c_void_p, C.c_void_p]),
    "mfix_under_relax_s_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mfix_calc_resid_c": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "mfix_calc_resid_c_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                        C.c_void_p]),
    "mfix_calc_resid_s": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                    C.c_double, C.c_int, C.c_void_p]),
    "mfix_calc_resid_s_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                        C.c_double, C.c_int, C.c_void_p, C.c_void_p]),
    "mfix_calc_resid_pp": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_void_p]),
    "mfix_calc_resid_pp_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_void_p, C.c_void_p]),
    "mfix_leq_resid_relax_bicgs_dev": (C.c_int, [C.c_int, C.c_void_p, C.c_double, C.c_void_p, C.c_double,
                                                 C.c_double, C.c_int, C.c_void_p, C.c_double, C.c_char_p,
                                                 C.c_double, C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mfix_leq_dev_alloc": (C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    "mfix_leq_dev_free": (C.c_int, [C.c_void_p]),
    "mfix_leq_copy_to_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "mfix_leq_copy_from_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "mfix_leq_copy_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "mfix_leq_sync": (C.c_int, []),
    "mfix_leq_copy_kernel_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    "mfix_leq_set_noise": (C.c_int, [C.c_void_p, C.c_uint64]),
    "mfix_leq_set_write_back": (C.c_int, [C.c_int]),
    "mfix_leq_last_stats": (C.c_int, [C.POINTER(Stats)]),
    "mfix_leq_prof_enable": (C.c_int, [C.c_int, C.c_int]),
    "mfix_leq_prof_read": (C.c_int, [C.c_int, C.POINTER(C.c_int64), c_double_p, c_double_p]),
    "mfix_leq_prof_idle": (C.c_int, [C.c_int, C.POINTER(C.c_int64), c_double_p]),
    "mfix_leq_msolve32_dev": (C.c_int, [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
}

EXPORTED = tuple(_SIGS)

_lib = None


def load(path: Path | str | None = None) -> C.CDLL:
    """Load libmfix_leq.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: run `python -m mfix_amd.build` (no fallback path exists)")
    lib = C.CDLL(str(p))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class LeqError(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = where
        try:
            s = load().mfix_leq_error_string(code)
            msg = f"{where}: {s.decode() if s else code}"
        except Exception:  # pragma: no cover - error path of the error path
            msg = f"{where}: code {code}"
        super().__init__(msg)
        self.code = code


def check(code: int, where: str) -> int:
    if code < 0:
        raise LeqError(code, where)
    return code
