"""Synthetic MFIX-layout systems (host side, deterministic per seed).

Thin numpy wrappers over ``mfix_leq_gen_system`` / ``mfix_leq_gen_uniform``
of libmfix_leq.so.  Arrays follow the reference's serial storage: A_M is
DIMENSION_3 x 7 column-major (seven contiguous planes l=-3..3, plane l at
offset (l+3)*N3; model/ambm_mod.f:27-28), J is the stride-1 axis
(model/include/functions.inc:37-42).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

PRESSURE, MOMENTUM, VOID_PRESSURE = 0, 1, 2
KINDS = {"pressure": PRESSURE, "momentum": MOMENTUM, "void_pressure": VOID_PRESSURE}


def dims(imax: int, jmax: int, kmax: int, no_k: bool = False) -> tuple[int, int, int, int]:
    ni, nj, nk = imax + 2, jmax + 2, (1 if no_k else kmax + 2)
    return ni, nj, nk, ni * nj * nk


def gen_system(kind, imax: int, jmax: int, kmax: int, seed: int, nspheres: int = -1):
    """Return (A_M (7*N3,), B_M (N3,)) float64 for a 3-D grid."""
    k = KINDS.get(kind, kind) if isinstance(kind, str) else int(kind)
    n3 = dims(imax, jmax, kmax)[3]
    a = np.empty(7 * n3, dtype=np.float64)
    b = np.empty(n3, dtype=np.float64)
    rc = _lib.load().mfix_leq_gen_system(k, imax, jmax, kmax, C.c_uint64(seed), nspheres,
                                         a.ctypes.data_as(_lib.c_double_p),
                                         b.ctypes.data_as(_lib.c_double_p))
    _lib.check(rc, "mfix_leq_gen_system")
    return a, b


def gen_uniform(n: int, seed: int, stream: int = 0) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    _lib.check(_lib.load().mfix_leq_gen_uniform(n, C.c_uint64(seed), C.c_uint64(stream),
                                               out.ctypes.data_as(_lib.c_double_p)),
               "mfix_leq_gen_uniform")
    return out


def config_system(cfg: int, sys_index: int = 0, n: int | None = None):
    """Systems of the benchmark configurations (seed = 1000*cfg + sys_index).

    cfg 0: 32^3 symmetric pressure-correction; cfg 1: 128^3 upwind momentum;
    cfg 2: 256^3 void-fraction pressure-correction.  ``n`` overrides the side.
    """
    side = {0: 32, 1: 128, 2: 256}[cfg] if n is None else n
    kind = {0: PRESSURE, 1: MOMENTUM, 2: VOID_PRESSURE}[cfg]
    return gen_system(kind, side, side, side, 1000 * cfg + sys_index)


def batch_config(nsys: int = 64, seed: int = 3000, lo: int = 16, hi: int = 128):
    """The batched workload: sides log-uniform in [lo, hi], half symmetric
    pressure-correction and half upwind momentum systems, DIAG or LINE per
    system, each with its own seed.  Returns a list of dicts (no arrays)."""
    rng = np.random.default_rng(seed)
    sides = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), nsys))).astype(int)
    kinds = np.array([PRESSURE, MOMENTUM] * ((nsys + 1) // 2))[:nsys]
    rng.shuffle(kinds)
    pcs = rng.choice(["DIAG", "LINE"], size=nsys)
    return [dict(side=int(sides[q]), kind=int(kinds[q]), pc=str(pcs[q]), seed=seed + q) for q in range(nsys)]


def serving_requests(rate: float, n: int, seed: int = 4000, median: float = 64.0, sigma: float = 0.5,
                     lo: int = 24, hi: int = 160):
    """A Poisson request stream: arrival times (s) at `rate` per second,
    cube sides lognormal (median, sigma) clipped to [lo, hi], 60 % upwind
    momentum / 40 % symmetric pressure-correction, TOL 1e-8 or 1e-10."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    sides = np.clip(np.rint(median * np.exp(sigma * rng.standard_normal(n))), lo, hi).astype(int)
    kinds = np.where(rng.uniform(size=n) < 0.6, MOMENTUM, PRESSURE)
    tols = np.where(rng.uniform(size=n) < 0.5, 1e-8, 1e-10)
    return [dict(t=float(t[q]), side=int(sides[q]), kind=int(kinds[q]), tol=float(tols[q]), seed=seed + q)
            for q in range(n)]


def pp_g_fields(imax: int, jmax: int, kmax: int, seed: int) -> tuple[dict, np.ndarray]:
    """Seeded fields of a gas pressure-correction assembly on a closed box
    with a pressure-outflow face: face densities ROP_gE/N/T, uniform 1 cm
    cells, ROP_g near 1 with a small change since the last step, velocities
    in [-1,1], a small mass source, D_E/D_N/D_T in [0.1,1]; FLAG 100 on the
    ghost layer, 20 on the I- face (inflow), 11 on the I+ face (pressure
    outflow), 1 inside."""
    ni, nj, nk = imax + 2, jmax + 2, kmax + 2
    n3 = ni * nj * nk
    rng = np.random.default_rng(seed)
    dx = 0.01
    f = {
        "rop_ge": rng.uniform(0.5, 1.5, n3), "rop_gn": rng.uniform(0.5, 1.5, n3), "rop_gt": rng.uniform(0.5, 1.5, n3),
        "ayz": np.full(n3, dx * dx), "axz": np.full(n3, dx * dx), "axy": np.full(n3, dx * dx),
        "vol": np.full(n3, dx ** 3), "rop_g": rng.uniform(0.9, 1.1, n3), "u_g": rng.uniform(-1, 1, n3),
        "v_g": rng.uniform(-1, 1, n3), "w_g": rng.uniform(-1, 1, n3), "sum_r_g": rng.uniform(-1e-3, 1e-3, n3),
        "d_e": rng.uniform(0.1, 1.0, n3), "d_n": rng.uniform(0.1, 1.0, n3), "d_t": rng.uniform(0.1, 1.0, n3),
    }
    f["rop_go"] = f["rop_g"] * rng.uniform(0.99, 1.01, n3)
    flag = np.full((nk, ni, nj), 100, dtype=np.int32)
    flag[1:-1, 1:-1, 1:-1] = 1
    flag[1:-1, 0, 1:-1] = 20
    flag[1:-1, ni - 1, 1:-1] = 11
    return f, flag.reshape(-1)


def phi_fields(imax: int, jmax: int, kmax: int, seed: int, odt: float = 100.0) -> tuple[dict, np.ndarray]:
    """Seeded fields of a gas-phase scalar transport equation (the inputs of
    CONV_DIF_PHI, SOURCE_PHI and UNDER_RELAX_S) on a box of uniform 1 cm
    cells: DIF ~ 1e-5, velocities in [-1,1] with mass fluxes ROP*U*A, the
    accumulation source S_P = ROP_o*VOL*ODT and S_C = S_P*PHI_old, EP in
    [0.4, 1]; FLAG as in pp_g_fields (ghost walls, inflow I- face, pressure
    outflow I+ face, fluid inside).  The 1-D geometry is uniform (ODX_E =
    1/DX, FX_E = 1/2, OX = 1)."""
    ni, nj, nk = imax + 2, jmax + 2, kmax + 2
    n3 = ni * nj * nk
    rng = np.random.default_rng(seed)
    dx = 0.01
    u, v, w = (rng.uniform(-1, 1, n3) for _ in range(3))
    rop = rng.uniform(0.9, 1.1, n3)
    area = np.full(n3, dx * dx)
    s_p = rop * dx ** 3 * odt
    phi_old = rng.uniform(0.0, 1.0, n3)
    f = {
        "dif": rng.uniform(0.5e-5, 2e-5, n3), "u": u, "v": v, "w": w,
        "flux_e": rop * u * area, "flux_n": rop * v * area, "flux_t": rop * w * area,
        "ayz": area, "axz": area.copy(), "axy": area.copy(),
        "s_p": s_p, "s_c": s_p * phi_old, "ep": rng.uniform(0.4, 1.0, n3), "phi": phi_old.copy(),
    }
    for name, n in (("x", ni), ("y", nj), ("z", nk)):
        o = np.full(n, 1.0 / dx)
        fr = np.full(n, 0.5)
        f.update({{"x": "odx_e", "y": "ody_n", "z": "odz_t"}[name]: o,
                  {"x": "fx_e", "y": "fy_n", "z": "fz_t"}[name]: fr,
                  {"x": "fx_e_bar", "y": "fy_n_bar", "z": "fz_t_bar"}[name]: fr.copy()})
    f["ox"] = np.ones(ni)
    flag = np.full((nk, ni, nj), 100, dtype=np.int32)
    flag[1:-1, 1:-1, 1:-1] = 1
    flag[1:-1, 0, 1:-1] = 20
    flag[1:-1, ni - 1, 1:-1] = 11
    return f, flag.reshape(-1)
