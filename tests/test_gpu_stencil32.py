"""GPU numerics of the FP32 stencil row of the mixed-precision inner solve:
`mfix_leq_matvec32_dev` runs the solver's own first stencil pass
(`k_matvec_dots4`: four cells per thread with 16-byte accesses where the
grid allows, `Stencil<float>::row` per cell elsewhere).

The FP32 matrix is in difference form: y(c) = sum_l a_l * (x_nb(l) - x_c)
+ e_c * x_c, where e_c = A0 + sum_l a_l is the row excess held in plane 3.
The kernel forms each difference and product in FP32, sums them in FP64 in a
fixed order and rounds once to FP32.  Two checks on the same FP32 inputs:

* bit for bit against a numpy restatement of those operations (numpy's FP32
  array arithmetic rounds like the GPU's; the excess product of two FP32
  values is exact in FP64, so a fused multiply-add gives the same bits);
* against the oracle's FP64 LEQ_MATVEC (model/leqsol_mod.f:103-265) of the
  equivalent reference-layout row, A0 = e_c - sum_l a_l, within the FP32
  rounding of the six products and of the result - a wrong neighbour, a
  dropped term or a wrong plane fails this one.
"""
import numpy as np
import pytest

from tests.cases import GRIDS, n3_of

pytestmark = pytest.mark.gpu


def _axis(n, cyc):
    """Neighbour indices along one axis (0-based, ghosts included): the
    outward neighbour of an array end is the cell itself, and on a cyclic
    axis the interior cells next to the ghosts wrap (set_increments.f:72-112)."""
    i = np.arange(n)
    end = (i == 0) | (i == n - 1)
    m = np.maximum(i - 1, 0)
    p = np.minimum(i + 1, n - 1)
    if cyc:
        m = np.where(~end & (m == 0), n - 2, m)
        p = np.where(~end & (p == n - 1), 1, p)
    return m, p


def _neighbours(im, jm, km, no_k, cyc):
    ni, nj = im + 2, jm + 2
    nk = 1 if no_k else km + 2
    k, i, j = np.meshgrid(np.arange(nk), np.arange(ni), np.arange(nj), indexing="ij")
    k, i, j = k.ravel(), i.ravel(), j.ravel()
    c = j + i * nj + k * nj * ni
    im_, ip_ = _axis(ni, cyc[0])
    jm_, jp_ = _axis(nj, cyc[1])
    nb = {
        "jm": jm_[j] + i * nj + k * nj * ni, "jp": jp_[j] + i * nj + k * nj * ni,
        "im": j + im_[i] * nj + k * nj * ni, "ip": j + ip_[i] * nj + k * nj * ni,
    }
    if no_k:
        nb["km"] = nb["kp"] = c
    else:
        km_, kp_ = _axis(nk, cyc[2])
        nb["km"] = j + i * nj + km_[k] * nj * ni
        nb["kp"] = j + i * nj + kp_[k] * nj * ni
    return c, nb


def _restated(a, x, nb, do_k):
    """The kernel's operations in numpy: FP32 differences and products, FP64
    sums in the kernel's order, one FP32 rounding."""
    f8 = np.float64
    t = {l: a[p] * (x[nb[l]] - x) for l, p in (("jm", 1), ("im", 2), ("ip", 4), ("jp", 5), ("km", 0), ("kp", 6))}
    v = t["jm"].astype(f8) + t["im"].astype(f8)
    v += t["ip"].astype(f8)
    v += t["jp"].astype(f8)
    if do_k:
        v += t["km"].astype(f8)
        v += t["kp"].astype(f8)
    v += a[3].astype(f8) * x.astype(f8)
    return v.astype(np.float32)


# The solver's FP32 stencil pass takes four consecutive cells per thread on
# 3-D grids without cyclic axes, off the first and last K planes: grids whose
# J rows and N3 are not multiples of four (groups straddling rows and planes,
# coefficient planes not 16-byte aligned), and one whose planes span several
# waves, next to the per-cell cases above.
STENCIL_GRIDS = GRIDS + [
    ("v_33x29x17", 2, 33, 29, 17, False, (0, 0, 0)),
    ("p_40x62x22", 0, 40, 62, 22, False, (0, 0, 0)),
    ("m_64x64x64", 1, 64, 64, 64, False, (0, 0, 0)),
]


@pytest.mark.parametrize("case", STENCIL_GRIDS, ids=[c[0] for c in STENCIL_GRIDS])
def test_fp32_stencil_row(gpu, oracle_mod, case):
    name, _kind, im, jm, km, no_k, cyc = case
    gpu.grid_init(im, jm, km, no_k=no_k, cyclic=tuple(bool(v) for v in cyc))
    og = oracle_mod.grid(im, jm, km, no_k=no_k, cyclic=cyc)
    n3 = n3_of(im, jm, km, no_k)
    rng = np.random.default_rng(4100 + len(name))
    a = rng.uniform(0.05, 1.0, (7, n3)).astype(np.float32)
    if no_k:
        a[0] = a[6] = 0.0
    a[3] = rng.uniform(-4.0, -0.5, n3).astype(np.float32)  # excess, A0-like magnitude
    x = rng.uniform(-1.0, 1.0, n3).astype(np.float32)
    x[rng.uniform(size=n3) < 0.05] = 0.0

    da = gpu.DeviceArray.from_host(np.ascontiguousarray(a.reshape(-1)))
    dx = gpu.DeviceArray.from_host(x)
    dy = gpu.DeviceArray.from_host(np.zeros(n3, dtype=np.float32))
    gpu.matvec_dev(da, dx, dy)
    y = dy.download()
    assert y.dtype == np.float32

    _c, nb = _neighbours(im, jm, km, no_k, cyc)
    np.testing.assert_array_equal(y, _restated(a, x, nb, not no_k))

    # the same row in the reference layout, evaluated by the FP64 oracle
    a64 = a.astype(np.float64)
    off = [0, 1, 2, 4, 5, 6]
    ref = a64.copy()
    ref[3] = a64[3] - a64[off].sum(axis=0)
    y64 = oracle_mod.matvec(og, ref.reshape(-1), x.astype(np.float64))
    scale = np.abs(a64[3] * x) + sum(np.abs(a64[p]) * (np.abs(x[nb[l]]) + np.abs(x))
                                     for l, p in (("jm", 1), ("im", 2), ("ip", 4), ("jp", 5), ("km", 0), ("kp", 6)))
    err = np.abs(y.astype(np.float64) - y64)
    assert np.all(err <= 4.0 * np.finfo(np.float32).eps * scale + 1e-30), (name, float((err / (scale + 1e-30)).max()))
