"""Build the native libraries of mfix_amd in-tree.

* ``libmfix_leq.so``: HIP kernels for gfx950 + the C ABI declared in
  ``include/mfix_leq.h`` (hipcc, one shared object, no torch dependency).
* ``libmfix_leq_fortran.so``: the ISO_C_BINDING drop-in ``LEQ_BICGS`` /
  ``SOLVE_LIN_EQ`` shim (flang), linked against ``libmfix_leq.so``.

Run ``python -m mfix_amd.build`` or call :func:`build`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
FSRC = PKG / "fortran"
LIB = PKG / "libmfix_leq.so"
FLIB = PKG / "libmfix_leq_fortran.so"

ARCH = os.environ.get("MFIX_AMD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fopenmp",
    # packed FP32 (v_pk_*_f32) stays off: plain scalar FP32 code only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", str(ROOT / "include"),
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the MI355X build needs ROCm's hipcc")


def _flang() -> str | None:
    for cand in (os.environ.get("FLANG"), shutil.which("flang"), "/opt/rocm/bin/amdflang",
                 "/opt/rocm/lib/llvm/bin/flang"):
        if cand and Path(cand).exists():
            return cand
    return None


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd[:3])} ...")


def build_native(verbose: bool = False) -> Path:
    """Compile every HIP/C++ source under csrc/ into libmfix_leq.so."""
    hipcc = _hipcc()
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    objs = []
    for src in srcs:
        obj = objdir / (src.stem + ".o")
        cmd = [hipcc, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
        if src.suffix == ".cpp":
            cmd = [hipcc, "-x", "c++", *[f for f in HIP_FLAGS if not f.startswith("--offload")],
                   "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fopenmp", "-o", str(tmp), *objs])
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose: bool = False) -> Path | None:
    """Compile the Fortran drop-in shim (skipped when no flang is installed)."""
    fc = _flang()
    srcs = sorted(FSRC.glob("*.f90"))
    if fc is None or not srcs:
        return None
    objdir = PKG / "_build" / "fortran"
    objdir.mkdir(parents=True, exist_ok=True)
    objs = []
    for src in srcs:
        obj = objdir / (src.stem + ".o")
        cmd = [fc, "-cpp", "-O2", "-fPIC", "-c", str(src), "-o", str(obj), "-module-dir", str(objdir)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(str(obj))
    tmp = FLIB.with_suffix(".so.tmp")
    _run([fc, "-shared", "-o", str(tmp), *objs, f"-L{PKG}", "-lmfix_leq", f"-Wl,-rpath,{PKG}",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FLIB)
    return FLIB


def build(verbose: bool = False) -> None:
    build_native(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
