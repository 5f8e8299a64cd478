"""GPU parity of the path's kernels against the oracle (through the C ABI).

LEQ_MATVEC, LEQ_MSOLVE1 (DIAG), LEQ_MSOLVE (LINE; every sweep mode) and the
unit scaling are element-wise or line-sequential computations that the HIP
kernels evaluate in the reference's operation order without contraction, so
they must match the oracle bit for bit.  dot_product_par is a reduction whose
order differs (tree vs. sequential); it is checked against an exact sum.
"""
import math

import numpy as np
import pytest

from tests.cases import GRIDS, make_system, n3_of, rng_vec

pytestmark = pytest.mark.gpu


def _setup(gpu, oracle_mod, kind, im, jm, km, no_k, cyc, **kw):
    gpu.grid_init(im, jm, km, no_k=no_k, cyclic=tuple(bool(c) for c in cyc), **kw)
    og = oracle_mod.grid(im, jm, km, no_k=no_k, cyclic=cyc, **kw)
    a, b = make_system(kind, im, jm, km, no_k)
    return og, a, b


@pytest.mark.parametrize("case", GRIDS, ids=[c[0] for c in GRIDS])
def test_matvec_bit_exact(gpu, oracle_mod, case):
    name, kind, im, jm, km, no_k, cyc = case
    og, a, b = _setup(gpu, oracle_mod, kind, im, jm, km, no_k, cyc)
    x = rng_vec(a.size // 7, 3)
    da, dx = gpu.DeviceArray.from_host(a), gpu.DeviceArray.from_host(x)
    dy = gpu.DeviceArray(x.size)
    gpu.matvec_dev(da, dx, dy)
    np.testing.assert_array_equal(dy.download(), oracle_mod.matvec(og, a, x))


@pytest.mark.parametrize("case", GRIDS, ids=[c[0] for c in GRIDS])
def test_unit_scaling_bit_exact(gpu, oracle_mod, case):
    name, kind, im, jm, km, no_k, cyc = case
    og, a, b = _setup(gpu, oracle_mod, kind, im, jm, km, no_k, cyc)
    a_ref, b_ref, ier_ref = oracle_mod.unit_scale(og, a, b)
    da, db = gpu.DeviceArray.from_host(a), gpu.DeviceArray.from_host(b)
    assert gpu.scale_dev(da, db) == ier_ref == 0
    np.testing.assert_array_equal(da.download(), a_ref)
    np.testing.assert_array_equal(db.download(), b_ref)


def test_unit_scaling_zero_row(gpu, oracle_mod):
    og, a, b = _setup(gpu, oracle_mod, 0, 5, 4, 3, False, (0, 0, 0))
    n3 = b.size
    a.reshape(7, n3)[:, 17] = 0.0
    _, _, ier_ref = oracle_mod.unit_scale(og, a, b)
    da, db = gpu.DeviceArray.from_host(a), gpu.DeviceArray.from_host(b)
    assert ier_ref == -2
    assert gpu.scale_dev(da, db) == -2


@pytest.mark.parametrize("case", GRIDS, ids=[c[0] for c in GRIDS])
def test_diag_bit_exact(gpu, oracle_mod, case):
    name, kind, im, jm, km, no_k, cyc = case
    og, a, b = _setup(gpu, oracle_mod, kind, im, jm, km, no_k, cyc)
    r = rng_vec(b.size, 5)
    da, dr = gpu.DeviceArray.from_host(a), gpu.DeviceArray.from_host(r)
    dz = gpu.DeviceArray(r.size)
    gpu.msolve_dev("DIAG", "RSRS", da, dr, dz)
    np.testing.assert_array_equal(dz.download(), oracle_mod.msolve(og, "DIAG", a, r)[0])


LINE_CASES = [(c, m) for c in GRIDS for m in ("RSRS", "ISIS", "ASAS", "KSKS", "RIRS", "JSJS")]


@pytest.mark.parametrize("case,cmethod", LINE_CASES, ids=[f"{c[0]}-{m}" for c, m in LINE_CASES])
def test_line_msolve_bit_exact(gpu, oracle_mod, case, cmethod):
    name, kind, im, jm, km, no_k, cyc = case
    og, a, b = _setup(gpu, oracle_mod, kind, im, jm, km, no_k, cyc)
    # the solver applies LINE to the unit-scaled system
    a, b, _ = oracle_mod.unit_scale(og, a, b)
    r = rng_vec(b.size, 11)
    z_ref, info = oracle_mod.msolve(og, "LINE", a, r, cmethod)
    assert info == 0
    da, dr = gpu.DeviceArray.from_host(a), gpu.DeviceArray.from_host(r)
    dz = gpu.DeviceArray(r.size)
    gpu.msolve_dev("LINE", cmethod, da, dr, dz)
    np.testing.assert_array_equal(dz.download(), z_ref)


def test_line_msolve_pivoting_lines(gpu, oracle_mod):
    """Lines that are not diagonally dominant take DGTSV's row interchanges."""
    og, a, b = _setup(gpu, oracle_mod, 1, 7, 9, 5, False, (0, 0, 0))
    n3 = b.size
    a2 = a.reshape(7, n3)
    rng = np.random.default_rng(1)
    # inflate the J couplings on a third of the cells so |D| < |DL| somewhere
    sel = rng.random(n3) < 0.33
    a2[1, sel] *= 25.0
    a2[5, sel] *= 25.0
    r = rng_vec(n3, 2)
    for cm in ("RSRS", "ASAS"):
        z_ref, info = oracle_mod.msolve(og, "LINE", a, r, cm)
        da, dr = gpu.DeviceArray.from_host(a), gpu.DeviceArray.from_host(r)
        dz = gpu.DeviceArray(n3)
        gpu.msolve_dev("LINE", cm, da, dr, dz)
        np.testing.assert_array_equal(dz.download(), z_ref)


@pytest.mark.parametrize("serial", [True, False])
@pytest.mark.parametrize("case", GRIDS[:3] + GRIDS[5:], ids=[c[0] for c in GRIDS[:3] + GRIDS[5:]])
def test_dot_product_par(gpu, oracle_mod, case, serial):
    name, kind, im, jm, km, no_k, cyc = case
    og, a, b = _setup(gpu, oracle_mod, kind, im, jm, km, no_k, cyc, is_serial=serial)
    x, y = rng_vec(b.size, 1), rng_vec(b.size, 2)
    dx, dy = gpu.DeviceArray.from_host(x), gpu.DeviceArray.from_host(y)
    got = gpu.dot_dev(dx, dy)
    ref = oracle_mod.dot(og, x, y)
    # exact value of the same domain, for a tolerance independent of order
    if serial:
        exact = math.fsum(x * y)
    else:
        ni, nj, nk = im + 2, jm + 2, (1 if no_k else km + 2)
        m = np.zeros((nk, ni, nj), bool)
        if no_k:
            m[0, 1:-1, 1:-1] = True
        else:
            m[1:-1, 1:-1, 1:-1] = True
        exact = math.fsum((x * y)[m.reshape(-1)])
    bound = 1e-14 * np.sum(np.abs(x * y))
    assert abs(got - exact) <= bound
    assert abs(ref - exact) <= bound


# J lines longer than one 16-row tile of the red-black kernel (nj >= 33:
# several tiles, and a partial last tile), in 3-D, 2-D and with a cyclic axis
TILED = [
    ("m_11x45x9", 1, 11, 45, 9, False, (0, 0, 0)),
    ("p_14x34x7_cycz", 0, 14, 34, 7, False, (0, 0, 1)),
    ("m_9x127x6", 1, 9, 127, 6, False, (0, 0, 0)),
    ("m_20x50_2d", 1, 20, 50, 1, True, (0, 0, 0)),
]
TILED_CASES = [(c, m) for c in TILED for m in ("RSRS", "RSSS")]


@pytest.mark.parametrize("case,cmethod", TILED_CASES, ids=[f"{c[0]}-{m}" for c, m in TILED_CASES])
def test_line_msolve_multi_tile_bit_exact(gpu, oracle_mod, case, cmethod):
    """R/S sweep strings run the solver's red-black kernel (the one every
    LINE solve uses), whose lines here span several 16-row tiles."""
    name, kind, im, jm, km, no_k, cyc = case
    print(f"multi-tile line solve {name} {cmethod}", flush=True)
    og, a, b = _setup(gpu, oracle_mod, kind, im, jm, km, no_k, cyc)
    a, b, _ = oracle_mod.unit_scale(og, a, b)
    r = rng_vec(b.size, 13)
    z_ref, info = oracle_mod.msolve(og, "LINE", a, r, cmethod)
    assert info == 0
    da, dr = gpu.DeviceArray.from_host(a), gpu.DeviceArray.from_host(r)
    dz = gpu.DeviceArray(r.size)
    gpu.msolve_dev("LINE", cmethod, da, dr, dz)
    np.testing.assert_array_equal(dz.download(), z_ref)


def test_line_msolve_multi_tile_pivoting(gpu, oracle_mod):
    """Row interchanges on lines of several tiles, partial last tile."""
    og, a, b = _setup(gpu, oracle_mod, 1, 8, 40, 6, False, (0, 0, 0))
    n3 = b.size
    a2 = a.reshape(7, n3)
    sel = np.random.default_rng(3).random(n3) < 0.3
    a2[1, sel] *= 25.0
    a2[5, sel] *= 25.0
    r = rng_vec(n3, 4)
    z_ref, info = oracle_mod.msolve(og, "LINE", a, r, "RSRS")
    da, dr = gpu.DeviceArray.from_host(a), gpu.DeviceArray.from_host(r)
    dz = gpu.DeviceArray(n3)
    gpu.msolve_dev("LINE", "RSRS", da, dr, dz)
    np.testing.assert_array_equal(dz.download(), z_ref)


@pytest.mark.parametrize("serial", [True, False])
def test_dot_product_multi_block_finish(gpu, oracle_mod, serial):
    """170^3 (5,088,448 cells): more than 16,384 partials per reduction, so
    the finish runs in several blocks (level-2 sums, release increment,
    acquire in the last block); below that one block sums every partial."""
    n = 170
    gpu.grid_init(n, n, n, is_serial=serial)
    nn = n + 2
    x, y = rng_vec(nn ** 3, 21), rng_vec(nn ** 3, 22)
    dx, dy = gpu.DeviceArray.from_host(x), gpu.DeviceArray.from_host(y)
    for rep in range(3):  # the finish counter is reset by each launch's last block
        got = gpu.dot_dev(dx, dy)
        if serial:
            exact = math.fsum(x * y)
        else:
            m = np.zeros((nn, nn, nn), bool)
            m[1:-1, 1:-1, 1:-1] = True
            exact = math.fsum((x * y)[m.reshape(-1)])
        assert abs(got - exact) <= 1e-14 * np.sum(np.abs(x * y)), (rep, got, exact)


# Lines per block of the red-black kernel: every variant gives the
# reference's bits, on lines of several tiles (nj up to 129) and grids whose
# colour has more lines than one block of each variant holds.
LPB_CASES = [(c, lpb) for c in TILED + [("p_33x40x29", 0, 33, 40, 29, False, (0, 0, 0))] for lpb in (64, 32, 16)]


@pytest.mark.parametrize("case,lpb", LPB_CASES, ids=[f"{c[0]}-lpb{l}" for c, l in LPB_CASES])
def test_line_msolve_lines_per_block_bit_exact(gpu, oracle_mod, case, lpb):
    name, kind, im, jm, km, no_k, cyc = case
    print(f"line solve {name} lines per block {lpb}", flush=True)
    og, a, b = _setup(gpu, oracle_mod, kind, im, jm, km, no_k, cyc)
    a, b, _ = oracle_mod.unit_scale(og, a, b)
    r = rng_vec(b.size, 17)
    z_ref, info = oracle_mod.msolve(og, "LINE", a, r, "RSRS")
    assert info == 0
    da, dr = gpu.DeviceArray.from_host(a), gpu.DeviceArray.from_host(r)
    dz = gpu.DeviceArray(r.size)
    gpu.set_line_block(im, jm, km if not no_k else 1, lpb)
    try:
        gpu.msolve_dev("LINE", "RSRS", da, dr, dz)
    finally:
        gpu.set_line_block(im, jm, km if not no_k else 1, 0)
    np.testing.assert_array_equal(dz.download(), z_ref)
