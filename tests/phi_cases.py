"""Seeded scalar-transport inputs for INIT_AB_M + CONV_DIF_PHI (first-order
upwinding), SOURCE_PHI and UNDER_RELAX_S (model/conv_dif_phi.f,
model/source_phi.f, model/under_relax.f).

Each case: a test grid, a FLAG array (ghost walls, an inflow face, a
pressure-outflow face, scattered interior walls and flow cells), the
DIMENSION_3 fields of CONV_DIF_PHI with exact zeros in the face velocities
and the diffusion coefficient (the `>= ZERO` branches and the SMALL_NUMBER
floor of the harmonic mean), a non-uniform 1-D geometry, the SOURCE_PHI
inputs with dilute cells (EP <= DIL_EP_s, some exactly equal) and zero PHI,
a B_M for SOURCE_PHI's two branches, and an under-relaxation factor.
"""
from __future__ import annotations

import hashlib

import numpy as np

import oracle
from tests.cases import GRIDS
from tests.pp_cases import pp_flags

DIL_EP_S = 1.0e-4   # model/toleranc_mod.f:24


def phi_flags(im, jm, km, no_k, cyc, seed):
    """pp_flags' layout; a cyclic axis's two ghost layers carry FLAG 106
    (CYCLIC_AT), as SET_FLAGS gives them (model/set_flags.f:145-146)."""
    f = pp_flags(im, jm, km, no_k, seed)
    ni, nj = im + 2, jm + 2
    nk = 1 if no_k else km + 2
    f = f.reshape(nk, ni, nj)
    if cyc[0]:
        f[:, [0, ni - 1], :] = 106
    if cyc[1]:
        f[:, :, [0, nj - 1]] = 106
    if cyc[2] and not no_k:
        f[[0, nk - 1], :, :] = 106
    return f.reshape(-1)


def _zero_some(rng, x, frac):
    x = x.copy()
    x[rng.uniform(size=x.size) < frac] = 0.0
    return x


def phi_geometry(ni, nj, nk, rng, no_k):
    def axis(n):
        dx = 0.01 * rng.uniform(0.8, 1.25, n)
        dxe = 0.5 * (dx + np.roll(dx, -1))
        fx = rng.uniform(0.3, 0.7, n)
        return 1.0 / dxe, fx, 1.0 - fx
    odx_e, fx_e, fx_e_bar = axis(ni)
    ody_n, fy_n, fy_n_bar = axis(nj)
    odz_t, fz_t, fz_t_bar = axis(nk)
    ox = rng.uniform(0.5, 2.0, ni)
    return {"odx_e": odx_e, "ox": ox, "fx_e": fx_e, "fx_e_bar": fx_e_bar, "ody_n": ody_n, "fy_n": fy_n,
            "fy_n_bar": fy_n_bar, "odz_t": odz_t, "fz_t": fz_t, "fz_t_bar": fz_t_bar}


def phi_fields(n3, ni, nj, nk, no_k, seed):
    rng = np.random.default_rng(seed)
    dx2 = 1.0e-4
    u = _zero_some(rng, rng.uniform(-1, 1, n3), 0.05)
    v = _zero_some(rng, rng.uniform(-1, 1, n3), 0.05)
    w = _zero_some(rng, rng.uniform(-1, 1, n3), 0.05)
    f = {
        "dif": _zero_some(rng, rng.uniform(0.5e-3, 2e-3, n3), 0.03),
        "u": u, "v": v, "w": w,
        # mass fluxes ROP*U*A: the sign of the velocity, except at a few faces
        "flux_e": u * rng.uniform(0.5, 1.5, n3) * dx2, "flux_n": v * rng.uniform(0.5, 1.5, n3) * dx2,
        "flux_t": w * rng.uniform(0.5, 1.5, n3) * dx2,
        "ayz": dx2 * rng.uniform(0.9, 1.1, n3), "axz": dx2 * rng.uniform(0.9, 1.1, n3),
        "axy": dx2 * rng.uniform(0.9, 1.1, n3),
    }
    flip = rng.uniform(size=n3) < 0.03
    f["flux_e"][flip] *= -1.0
    ep = rng.uniform(0.3, 1.0, n3)
    u_ep = rng.uniform(size=n3)
    ep[u_ep < 0.06] = rng.uniform(0.0, DIL_EP_S, int((u_ep < 0.06).sum()))
    ep[(u_ep >= 0.06) & (u_ep < 0.08)] = DIL_EP_S
    f["s_p"] = rng.uniform(0.0, 1e-3, n3)
    f["s_c"] = rng.uniform(-1e-3, 1e-3, n3)
    f["ep"] = ep
    f["phi"] = _zero_some(rng, rng.uniform(-1, 1, n3), 0.05)
    f.update(phi_geometry(ni, nj, nk, rng, no_k))
    b_in = rng.uniform(-2e-3, 2e-3, n3)
    return f, b_in


def phi_cases():
    out = []
    for gi, (name, kind, im, jm, km, no_k, cyc) in enumerate(GRIDS):
        g = oracle.grid(im, jm, km, no_k=no_k, cyclic=cyc)
        n3 = oracle.n3_of(g)
        ni, nj = im + 2, jm + 2
        nk = 1 if no_k else km + 2
        flag = phi_flags(im, jm, km, no_k, cyc, 1200 + gi)
        fields, b_in = phi_fields(n3, ni, nj, nk, no_k, 1250 + gi)
        out.append(dict(name=f"{name}/ur0.5", grid=g, flag=flag, fields=fields, b_in=b_in, ur=0.5))
        if gi in (0, 4):
            out.append(dict(name=f"{name}/ur1", grid=g, flag=flag, fields=fields, b_in=b_in, ur=1.0))
    return out


def phi_input_sha(case) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(case["flag"]).tobytes())
    f = case["fields"]
    for k in oracle.PHI_FIELDS + oracle.PHI_SOURCES + oracle.PHI_GEO_I + oracle.PHI_GEO_J + oracle.PHI_GEO_K:
        h.update(np.ascontiguousarray(f[k]).tobytes())
    h.update(np.ascontiguousarray(case["b_in"]).tobytes())
    h.update(repr(case["ur"]).encode())
    return h.hexdigest()


def reference_chain(case, run):
    """The three stages as a chain: CONV_DIF_PHI from nothing, SOURCE_PHI on
    its A_M with b_in as B_M, UNDER_RELAX_S on SOURCE_PHI's output.  run is
    (stage, a, b) -> (a, b)."""
    a1, b1 = run(1, None, None)
    a2, b2 = run(2, a1, case["b_in"])
    a3, b3 = run(3, a2, b2)
    return (a1, b1), (a2, b2), (a3, b3)
