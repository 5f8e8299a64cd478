"""mfix_amd: MI355X-native replacement of MFIX's LEQ_BICGS linear-solver path.

The product is the C ABI library ``libmfix_leq.so`` (HIP kernels for
gfx950, declared in ``include/mfix_leq.h``) and the Fortran drop-in shim
``libmfix_leq_fortran.so``.  This package builds them (``mfix_amd.build``)
and mirrors the reference's LEQ interface for Python callers
(``mfix_amd.leq``).
"""
from __future__ import annotations

from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libmfix_leq.so"
FORTRAN_LIB_PATH = PKG_DIR / "libmfix_leq_fortran.so"

__all__ = ["PKG_DIR", "LIB_PATH", "FORTRAN_LIB_PATH"]
