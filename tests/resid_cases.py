"""Seeded residual cases (CALC_RESID_C / _S / _PP, model/calc_resid.f).

Each case: a test grid and system (tests.cases), a FLAG array (MFIX cell
types: 1 fluid, 10-31 flow boundaries, >= 100 walls; model/include/
functions.inc:409-474) with ghost walls, one inflow face and scattered
interior wall and flow cells, a VAR vector, and the routine's scalars.
"""
from __future__ import annotations

import hashlib

import numpy as np

import oracle
from tests.cases import GRIDS, make_system


def flags_for(im, jm, km, no_k, seed, pattern="mixed"):
    ni, nj = im + 2, jm + 2
    nk = 1 if no_k else km + 2
    rng = np.random.default_rng(seed)
    f = np.full((nk, ni, nj), 100, dtype=np.int32)          # ghost cells: walls
    ks = slice(0, 1) if no_k else slice(1, nk - 1)
    f[ks, 1:ni - 1, 1:nj - 1] = 1                            # interior: fluid
    if pattern == "allwall":
        f[:] = 100
    elif pattern == "mixed":
        f[ks, 0, 1:nj - 1] = 20                              # an inflow face (I = 1 ghosts)
        inner = f[ks, 1:ni - 1, 1:nj - 1]
        u = rng.uniform(size=inner.shape)
        inner[u < 0.08] = 100                                # interior walls
        inner[(u >= 0.08) & (u < 0.12)] = 10                 # interior flow cells
        inner[(u >= 0.12) & (u < 0.14)] = 31
    return f.reshape(-1)


def resid_cases():
    out = []
    for gi, (name, kind, im, jm, km, no_k, cyc) in enumerate(GRIDS):
        g = oracle.grid(im, jm, km, no_k=no_k, cyclic=cyc)
        a, b = make_system(kind, im, jm, km, no_k)
        n3 = b.size
        rng = np.random.default_rng(7000 + gi)
        x = rng.uniform(-1, 1, n3)
        flag = flags_for(im, jm, km, no_k, 500 + gi)
        ep_s = rng.uniform(0.0, 2.0e-4, n3)
        for debug in (True, False):
            d = int(debug)
            out.append(dict(name=f"{name}/C/{d}", grid=g, kind="C", a=a, b=b, x=x, flag=flag, debug=debug, m=0,
                            tol=0.0, norm=1.0, ep_s=None))
            out.append(dict(name=f"{name}/S/{d}", grid=g, kind="S", a=a, b=b, x=x, flag=flag, debug=debug, m=0,
                            tol=0.3, norm=1.0, ep_s=None))
            out.append(dict(name=f"{name}/S1/{d}", grid=g, kind="S", a=a, b=b, x=x, flag=flag, debug=debug, m=1,
                            tol=0.05, norm=1.0, ep_s=ep_s))
            out.append(dict(name=f"{name}/PP/{d}", grid=g, kind="PP", a=a, b=b, x=x, flag=flag, debug=debug, m=0,
                            tol=0.0, norm=0.37, ep_s=None))
    # special cases on the first grid
    name, kind, im, jm, km, no_k, cyc = GRIDS[0]
    g = oracle.grid(im, jm, km, no_k=no_k, cyclic=cyc)
    a, b = make_system(kind, im, jm, km, no_k)
    n3 = b.size
    x = np.random.default_rng(11).uniform(-1, 1, n3)
    wall = flags_for(im, jm, km, no_k, 1, pattern="allwall")
    fluid = flags_for(im, jm, km, no_k, 1, pattern="fluid")
    for debug in (True, False):
        d = int(debug)
        for kd in ("C", "S", "PP"):
            out.append(dict(name=f"special/nofluid/{kd}/{d}", grid=g, kind=kd, a=a, b=b, x=x, flag=wall, debug=debug,
                            m=0, tol=0.0, norm=1.0, ep_s=None))
        # VAR = 0: DEN = 0 while NUM > 0 (RESID = UNDEFINED)
        out.append(dict(name=f"special/zerovar/C/{d}", grid=g, kind="C", a=a, b=b, x=np.zeros(n3), flag=fluid,
                        debug=debug, m=0, tol=0.0, norm=1.0, ep_s=None))
        # B = 0 for the pressure correction; NORM = 0 (RESID = LARGE_NUMBER)
        out.append(dict(name=f"special/zerob/PP/{d}", grid=g, kind="PP", a=a, b=np.zeros(n3), x=x, flag=fluid,
                        debug=debug, m=0, tol=0.0, norm=1.0, ep_s=None))
        out.append(dict(name=f"special/zeronorm/PP/{d}", grid=g, kind="PP", a=a, b=b, x=x, flag=fluid,
                        debug=debug, m=0, tol=0.0, norm=0.0, ep_s=None))
        # a system solved exactly has NUM = 0 only up to rounding; the plain
        # fluid box checks the wall-free neighbour path
        out.append(dict(name=f"special/fluid/C/{d}", grid=g, kind="C", a=a, b=b, x=x, flag=fluid, debug=debug, m=0,
                        tol=0.0, norm=1.0, ep_s=None))
    return out


def input_sha(case) -> str:
    h = hashlib.sha256()
    for k in ("a", "b", "x", "flag", "ep_s"):
        v = case[k]
        if v is not None:
            h.update(np.ascontiguousarray(v).tobytes())
    h.update(repr((case["kind"], case["debug"], case["m"], case["tol"], case["norm"])).encode())
    return h.hexdigest()
