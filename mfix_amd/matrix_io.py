"""Readers of MFIX's two matrix dump formats, so that systems an MFIX run
wrote out can be replayed through the device solver.

* matrixExtract (model/dbg_mod.f:228-552, after initExtract, :82-190):
  Am[_VAR][_PASS].csv holds one line per cell of the extraction region, K
  outermost, then I, then J (dbg_mod.f:310-312), with the row's coefficients
  A_M(ijk,-3:3) in 3-D or A_M(ijk,-2:2) in 2-D as `e14.6` values separated
  by commas (:496-502); a coefficient whose neighbour lies outside the region
  is written as 0 (:398-470).  Bm[...].csv holds B_M(ijk), one `e14.6` per
  line.  The region defaults to the whole domain IMIN3..IMAX3, JMIN3..JMAX3,
  KMIN3..KMAX3 (:119-124).
* WRITE_AB_M (model/write_ab_m.f:20-145) writes, after six header lines, one
  log line per cell of IMIN2..IMAX2 x JMIN2..JMAX2 x KMIN2..KMAX2 in the same
  K, I, J order, as `(I5, 3(I3), 8(1X,G9.2))`: the I-fastest cell number,
  I, J, K, A_M(ijk,-3:3) and B_M(ijk) (:117-119).  G9.2 keeps two
  significant digits; I5 overflows to asterisks past 99999 cells, so the
  reader places each line by its I, J, K columns.

Both readers return A_M (7 planes) and B_M in this library's storage
contract (the reference's serial layout, one ghost layer: IMIN3 = 1,
IMAX3 = IMAX + 2), ready for mfix_leq_bicgs / bicgs_dev.  The dumps carry
the values at their text precision only (7 and 2 significant digits).
"""
from __future__ import annotations

import argparse
import json
from pathlib import Path

import numpy as np


def _dims(imax: int, jmax: int, kmax: int, no_k: bool):
    ni, nj = imax + 2, jmax + 2
    nk = 1 if no_k else kmax + 2
    return ni, nj, nk


def _read_csv(path, ncol: int | None = None) -> np.ndarray:
    try:
        import pandas as pd
        arr = pd.read_csv(path, header=None, dtype=np.float64, skipinitialspace=True).to_numpy()
    except ImportError:  # pragma: no cover - pandas is part of the image
        arr = np.loadtxt(path, delimiter=",", dtype=np.float64, ndmin=2)
    if ncol is not None and arr.shape[1] != ncol:
        raise ValueError(f"{path}: {arr.shape[1]} columns, expected {ncol}")
    return arr


def read_matrix_extract(am_path, bm_path, imax: int, jmax: int, kmax: int = 1, *, no_k: bool = False,
                        region: tuple[int, int, int, int, int, int] | None = None):
    """A_M, B_M of a matrixExtract dump of a grid with IMAX, JMAX, KMAX
    interior cells.  region = (iLB, iUB, jLB, jUB, kLB, kUB), 1-based MFIX
    indices, when initExtract was given bounds; cells outside it stay 0
    (A0 = 0 too: such rows are not a solvable system until filled)."""
    ni, nj, nk = _dims(imax, jmax, kmax, no_k)
    ilb, iub, jlb, jub, klb, kub = region or (1, ni, 1, nj, 1, nk)
    if not (1 <= ilb <= iub <= ni and 1 <= jlb <= jub <= nj and 1 <= klb <= kub <= nk):
        raise ValueError(f"extraction region {region} outside the grid {ni}x{nj}x{nk}")
    am = _read_csv(am_path)
    bm = _read_csv(bm_path, 1)[:, 0]
    planes = {7: range(-3, 4), 5: range(-2, 3)}.get(am.shape[1])
    if planes is None:
        raise ValueError(f"{am_path}: {am.shape[1]} coefficients per row (7 in 3-D, 5 in 2-D)")
    if (am.shape[1] == 5) != no_k:
        raise ValueError(f"{am_path}: {am.shape[1]} coefficients per row for a {'2' if no_k else '3'}-D grid")
    ri, rj, rk = iub - ilb + 1, jub - jlb + 1, kub - klb + 1
    if am.shape[0] != ri * rj * rk or bm.size != ri * rj * rk:
        raise ValueError(f"{am.shape[0]} A rows and {bm.size} B rows for a region of {ri * rj * rk} cells")
    n3 = ni * nj * nk
    # rows in K, I, J order; storage index (j-1) + (i-1)*nj + (k-1)*nj*ni
    k, i, j = np.meshgrid(np.arange(klb, kub + 1), np.arange(ilb, iub + 1), np.arange(jlb, jub + 1), indexing="ij")
    c = ((j - 1) + (i - 1) * nj + (k - 1) * nj * ni).ravel()
    a = np.zeros((7, n3))
    for col, l in enumerate(planes):
        a[l + 3, c] = am[:, col]
    b = np.zeros(n3)
    b[c] = bm
    return a.reshape(-1), b


def read_write_ab_m(log_path):
    """A_M, B_M and (imax, jmax, kmax, no_k) of a WRITE_AB_M log; the grid
    is the largest I, J, K the lines name (IMAX2 = IMAX + 2)."""
    rows = []
    with open(log_path) as f:
        for line in f:
            if len(line.rstrip("\n")) < 14 + 8 * 10:
                continue
            try:
                i, j, k = int(line[5:8]), int(line[8:11]), int(line[11:14])
                vals = [float(line[14 + 10 * q: 24 + 10 * q]) for q in range(8)]
            except ValueError:
                continue   # header lines, or an I3 field past 999
            rows.append((i, j, k, *vals))
    if not rows:
        raise ValueError(f"{log_path}: no WRITE_AB_M lines")
    t = np.asarray(rows)
    ijk = t[:, :3].astype(np.int64)
    ni, nj, nk = (int(v) for v in ijk.max(axis=0))
    if (ijk.min(axis=0) != 1).any() or t.shape[0] != ni * nj * nk:
        raise ValueError(f"{log_path}: {t.shape[0]} lines do not cover a {ni}x{nj}x{nk} grid from 1")
    no_k = nk == 1
    c = (ijk[:, 1] - 1) + (ijk[:, 0] - 1) * nj + (ijk[:, 2] - 1) * nj * ni
    n3 = ni * nj * nk
    a = np.zeros((7, n3))
    a[:, c] = t[:, 3:10].T
    b = np.zeros(n3)
    b[c] = t[:, 10]
    return a.reshape(-1), b, (ni - 2, nj - 2, 0 if no_k else nk - 2, no_k)


def replay(a_m: np.ndarray, b_m: np.ndarray, imax: int, jmax: int, kmax: int, *, no_k: bool = False,
           pc: str = "DIAG", cmethod: str = "RSRS", tol: float = 1e-10, itmax: int = 500, mixed: bool = False,
           var: np.ndarray | None = None) -> dict:
    """Solve a dumped system on the device (LEQ_BICGS semantics; mixed runs
    the FP32/FP64 solver).  Returns the solution, IER and the solve stats."""
    from . import leq
    leq.init(0)
    leq.grid_init(imax, jmax, kmax, no_k=no_k)
    n3 = b_m.size
    x = np.zeros(n3) if var is None else np.asarray(var, dtype=np.float64)
    dx, da, db = (leq.DeviceArray.from_host(np.ascontiguousarray(v, dtype=np.float64)) for v in (x, a_m, b_m))
    out = leq.bicgs_dev(dx, da, db, pc=pc, cmethod=cmethod, tol=tol, itmax=itmax, mixed=mixed)
    return {"x": dx.download(), "ier": out.ier, "stats": out.stats}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Replay an MFIX matrix dump through the device solver.")
    src = ap.add_mutually_exclusive_group(required=True)
    src.add_argument("--am", help="matrixExtract Am*.csv (needs --bm and --dims)")
    src.add_argument("--log", help="WRITE_AB_M log file")
    ap.add_argument("--bm")
    ap.add_argument("--dims", type=int, nargs=3, metavar=("IMAX", "JMAX", "KMAX"))
    ap.add_argument("--no-k", action="store_true")
    ap.add_argument("--pc", default="DIAG")
    ap.add_argument("--sweep", default="RSRS")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--itmax", type=int, default=500)
    ap.add_argument("--mixed", action="store_true")
    args = ap.parse_args(argv)
    if args.am:
        if not args.bm or not args.dims:
            ap.error("--am needs --bm and --dims")
        im, jm, km = args.dims
        a, b = read_matrix_extract(args.am, args.bm, im, jm, km, no_k=args.no_k)
        no_k = args.no_k
    else:
        a, b, (im, jm, km, no_k) = read_write_ab_m(args.log)
    r = replay(a, b, im, jm, max(km, 1), no_k=no_k, pc=args.pc, cmethod=args.sweep, tol=args.tol,
               itmax=args.itmax, mixed=args.mixed)
    print(json.dumps({"source": str(Path(args.am or args.log).name), "grid": [im, jm, km, no_k],
                      "ier": r["ier"], "stats": r["stats"]}))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
