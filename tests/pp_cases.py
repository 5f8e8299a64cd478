"""Seeded gas pressure-correction fields (INIT_AB_M + CONV_PP_G +
SOURCE_PP_G, model/conv_pp_g.f, model/source_pp_g.f).

Each case: a test grid, a FLAG array (ghost walls, an inflow face, a
pressure-outflow face (FLAG 11) and scattered interior walls), the sixteen
DIMENSION_3 fields the routines read, ODT and the reference cell IJK_P_g.
"""
from __future__ import annotations

import hashlib

import numpy as np

import oracle
from tests.cases import GRIDS


def pp_flags(im, jm, km, no_k, seed):
    ni, nj = im + 2, jm + 2
    nk = 1 if no_k else km + 2
    rng = np.random.default_rng(seed)
    f = np.full((nk, ni, nj), 100, dtype=np.int32)
    ks = slice(0, 1) if no_k else slice(1, nk - 1)
    f[ks, 1:ni - 1, 1:nj - 1] = 1
    f[ks, 0, 1:nj - 1] = 20                  # mass inflow face
    f[ks, ni - 1, 1:nj - 1] = 11             # pressure outflow face (P_FLOW_AT)
    inner = f[ks, 1:ni - 1, 1:nj - 1]
    u = rng.uniform(size=inner.shape)
    inner[u < 0.06] = 100
    inner[(u >= 0.06) & (u < 0.08)] = 10     # interior pressure-inflow cells
    return f.reshape(-1)


def pp_fields(n3, seed):
    rng = np.random.default_rng(seed)
    dx = 0.01
    f = {
        "rop_ge": rng.uniform(0.5, 1.5, n3), "rop_gn": rng.uniform(0.5, 1.5, n3), "rop_gt": rng.uniform(0.5, 1.5, n3),
        "ayz": np.full(n3, dx * dx) * rng.uniform(0.9, 1.1, n3), "axz": np.full(n3, dx * dx),
        "axy": np.full(n3, dx * dx) * rng.uniform(0.9, 1.1, n3), "vol": np.full(n3, dx ** 3),
        "rop_g": rng.uniform(0.9, 1.1, n3), "u_g": rng.uniform(-1, 1, n3), "v_g": rng.uniform(-1, 1, n3),
        "w_g": rng.uniform(-1, 1, n3), "sum_r_g": rng.uniform(-1e-3, 1e-3, n3),
        "d_e": rng.uniform(0.1, 1.0, n3), "d_n": rng.uniform(0.1, 1.0, n3), "d_t": rng.uniform(0.1, 1.0, n3),
    }
    f["rop_go"] = f["rop_g"] * rng.uniform(0.99, 1.01, n3)
    return f


def pp_cases():
    out = []
    for gi, (name, kind, im, jm, km, no_k, cyc) in enumerate(GRIDS):
        g = oracle.grid(im, jm, km, no_k=no_k, cyclic=cyc)
        n3 = oracle.n3_of(g)
        flag = pp_flags(im, jm, km, no_k, 900 + gi)
        fields = pp_fields(n3, 950 + gi)
        fluid = np.flatnonzero(flag == 1)
        out.append(dict(name=f"{name}/plain", grid=g, flag=flag, fields=fields, odt=100.0, ijk_p_g=0))
        out.append(dict(name=f"{name}/pref", grid=g, flag=flag, fields=fields, odt=250.0,
                        ijk_p_g=int(fluid[len(fluid) // 2]) + 1))
    # a fluid cell whose scaled faces all vanish and whose balance is zero:
    # A0 = 0 with B = 0 becomes A0 = -1, B = 0 (source_pp_g.f:166-170)
    name, kind, im, jm, km, no_k, cyc = GRIDS[0]
    g = oracle.grid(im, jm, km, no_k=no_k, cyclic=cyc)
    n3 = oracle.n3_of(g)
    flag = pp_flags(im, jm, km, no_k, 42)
    nj, ni = jm + 2, im + 2
    c = (1 + jm // 2) + (1 + im // 2) * nj + (1 + km // 2) * nj * ni
    flag[[c, c - 1, c - nj, c - nj * ni]] = 1
    fields = {k: v.copy() for k, v in pp_fields(n3, 43).items()}
    for key, nbr in (("d_e", (c, c - nj)), ("d_n", (c, c - 1)), ("d_t", (c, c - nj * ni))):
        fields[key][list(nbr)] = 0.0
    for key, nbr in (("u_g", (c, c - nj)), ("v_g", (c, c - 1)), ("w_g", (c, c - nj * ni))):
        fields[key][list(nbr)] = 0.0
    fields["rop_go"][c] = fields["rop_g"][c]
    fields["sum_r_g"][c] = 0.0
    out.append(dict(name="special/zeroA0", grid=g, flag=flag, fields=fields, odt=100.0, ijk_p_g=0))
    return out


def pp_input_sha(case) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(case["flag"]).tobytes())
    for k in oracle.PP_FIELDS:
        h.update(np.ascontiguousarray(case["fields"][k]).tobytes())
    h.update(repr((case["odt"], case["ijk_p_g"])).encode())
    return h.hexdigest()
