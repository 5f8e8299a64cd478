"""GPU: the LINE preconditioner of the mixed-precision inner solve
(k_line_pcr_sys, through mfix_leq_msolve32_dev) against the oracle.

The FP32 matrix is the one the mixed solver builds from the unit-scaled
system (couplings rounded to FP32, A0 rounded to FP32 in plane 7); the
oracle runs the reference's LEQ_MSOLVE (model/leqsol_mod.f:286-598, DGTSV
per line) in FP64 on exactly those rounded values and the same right-hand
side.  The several-lanes-per-line solve rounds differently, so the bar is
FP32 accuracy: max-norm difference <= 2e-5 of max |z|.  The grids give J
lines of every length class of the kernel's variants (idle lanes, exactly
full lanes, one row per lane), 2-D and cyclic axes.
"""
import numpy as np
import pytest

from tests.cases import make_system

pytestmark = pytest.mark.gpu

# (name, kind, imax, jmax, kmax, no_k, cyclic)
LINE32 = [
    ("j14", 1, 6, 14, 5, False, (0, 0, 0)),      # 16 rows: one row per lane
    ("j20", 1, 7, 20, 5, False, (1, 0, 0)),
    ("j30", 0, 5, 30, 6, False, (0, 0, 0)),      # 32 rows: lanes exactly full
    ("j50", 1, 6, 50, 5, False, (0, 0, 1)),
    ("j100", 1, 5, 100, 4, False, (0, 0, 0)),
    ("j126", 0, 4, 126, 3, False, (0, 0, 0)),    # 128 rows: 32 lanes of 4, full
    ("j128", 1, 4, 128, 4, False, (0, 0, 0)),    # the 128^3 line length
    ("j158", 1, 3, 158, 3, False, (1, 0, 0)),    # 160 rows: 32 lanes of 5, full
    ("j200", 0, 4, 200, 3, False, (0, 0, 0)),
    ("j254", 1, 3, 254, 3, False, (0, 0, 0)),    # 256 rows: 64 lanes of 4, full
    ("j256", 1, 3, 256, 3, False, (0, 0, 0)),    # the 256^3 line length
    ("j300", 1, 3, 300, 3, False, (0, 0, 0)),
    ("j450", 1, 3, 450, 3, False, (0, 0, 0)),
    ("j60_2d", 1, 30, 60, 1, True, (0, 0, 0)),
]


def _fp32_matrix(oracle_mod, og, a, b):
    """Unit-scaled system -> the inner solve's 8 FP32 planes, and the FP64
    matrix with the same (rounded) couplings and diagonal for the oracle."""
    a_s, b_s, _ = oracle_mod.unit_scale(og, a, b)
    n3 = b.size
    p = a_s.reshape(7, n3)
    a32 = np.zeros((8, n3), np.float32)
    for q in (0, 1, 2, 4, 5, 6):
        a32[q] = p[q].astype(np.float32)
    a32[3] = (p[3] + p[[0, 1, 2, 4, 5, 6]].sum(axis=0)).astype(np.float32)
    a32[7] = p[3].astype(np.float32)
    a64 = a32[:7].astype(np.float64)
    a64[3] = a32[7].astype(np.float64)
    return a32.reshape(-1), a64.reshape(-1)


@pytest.mark.parametrize("case", LINE32, ids=[c[0] for c in LINE32])
@pytest.mark.parametrize("cmethod", ["RSRS", "RSSS"])
def test_line32_against_oracle(gpu, oracle_mod, case, cmethod):
    name, kind, im, jm, km, no_k, cyc = case
    print(f"FP32 line solve {name} {cmethod}", flush=True)
    gpu.grid_init(im, jm, km, no_k=no_k, cyclic=tuple(bool(c) for c in cyc))
    og = oracle_mod.grid(im, jm, km, no_k=no_k, cyclic=cyc)
    a, b = make_system(kind, im, jm, km, no_k, seed=31)
    a32, a64 = _fp32_matrix(oracle_mod, og, a, b)
    r32 = np.random.default_rng(7).uniform(-1, 1, b.size).astype(np.float32)
    z_ref, info = oracle_mod.msolve(og, "LINE", a64, r32.astype(np.float64), cmethod)
    assert info == 0
    da = gpu.DeviceArray.from_host(a32)
    dr = gpu.DeviceArray.from_host(r32)
    dz = gpu.DeviceArray(b.size, np.float32)
    gpu.msolve32_dev(cmethod, da, dr, dz)
    z = dz.download().astype(np.float64)
    err = np.max(np.abs(z - z_ref)) / np.max(np.abs(z_ref))
    assert err <= 2e-5, err
