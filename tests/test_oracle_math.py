"""The oracle against what the mathematics fixes (no reference involved).

* LEQ_MATVEC vs an independently built scipy CSR matrix of the 7-point
  system (numpy index arithmetic for the clamped/cyclic neighbours);
* DGTSV vs LAPACK's dgtsv from scipy (same algorithm, partial pivoting) and
  vs a dense solve, incl. INFO on a zero pivot;
* DIAG: z * A0 == r; LINE: after a red-black sweep the lines of the last
  colour satisfy their tridiagonal equations;
* LEQ_BICGS: the solution meets the residual bound of the scaled system and
  agrees with a sparse direct solve to the conditioning bound; closed-form
  special cases (identity, B = 0, zero row, ITMAX classification).
Each check is chosen so a plausible mistake fails it: a planted stride swap
and a planted sign error are shown to be caught.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse.linalg as spla

from tests.cases import GRIDS, make_system, rng_vec, to_csr


@pytest.mark.parametrize("case", GRIDS, ids=[c[0] for c in GRIDS])
def test_matvec_equals_sparse_matrix(oracle_mod, case):
    name, kind, im, jm, km, no_k, cyc = case
    g = oracle_mod.grid(im, jm, km, no_k=no_k, cyclic=cyc)
    a, b = make_system(kind, im, jm, km, no_k)
    x = rng_vec(b.size, 3)
    y = oracle_mod.matvec(g, a, x)
    ref = to_csr(a, im, jm, km, no_k, cyc) @ x
    assert np.max(np.abs(y - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_planted_stride_swap_is_caught(oracle_mod):
    """Swapping the south (J-) and west (I-) planes must break the CSR check."""
    im, jm, km = 7, 6, 5
    g = oracle_mod.grid(im, jm, km)
    a, b = make_system(1, im, jm, km)
    x = rng_vec(b.size, 3)
    bad = a.reshape(7, -1).copy()
    bad[[1, 2]] = bad[[2, 1]]
    y_bad = oracle_mod.matvec(g, bad.reshape(-1), x)
    ref = to_csr(a, im, jm, km) @ x
    assert np.max(np.abs(y_bad - ref)) > 1e-3 * np.max(np.abs(ref))


def _random_tridiag(n, rng, dominant=True):
    dl = rng.uniform(-1, 1, n - 1)
    du = rng.uniform(-1, 1, n - 1)
    d = rng.uniform(-1, 1, n)
    if dominant:
        d = np.sign(d) * (2.5 + np.abs(d))
    return dl, d, du


@pytest.mark.parametrize("dominant", [True, False])
@pytest.mark.parametrize("n", [1, 2, 3, 7, 34, 130])
def test_dgtsv_matches_lapack(oracle_mod, n, dominant):
    rng = np.random.default_rng(n + 100 * dominant)
    for _ in range(5):
        dl, d, du = _random_tridiag(n, rng, dominant)
        b = rng.uniform(-1, 1, n)
        x, info = oracle_mod.dgtsv(dl, d, du, b)
        if n == 1:  # scipy's wrapper rejects the empty off-diagonals
            assert info == 0 and x[0] == b[0] / d[0]
            continue
        _, _, _, xl, infol = sla.lapack.dgtsv(dl.copy(), d.copy(), du.copy(), b.copy())
        assert info == infol == 0
        # same algorithm: equal to rounding (LAPACK builds may contract a*b+c)
        np.testing.assert_allclose(x, xl, rtol=1e-12, atol=1e-14)
        dense = np.diag(d) + np.diag(dl, -1) + np.diag(du, 1)
        np.testing.assert_allclose(dense @ x, b, rtol=1e-10, atol=1e-12)


def test_dgtsv_zero_pivot_info(oracle_mod):
    d = np.array([0.0, 1.0, 1.0])
    x, info = oracle_mod.dgtsv([0.0, 0.0], d, [1.0, 0.0], [1.0, 1.0, 1.0])
    _, _, _, _, infol = sla.lapack.dgtsv(np.zeros(2), d.copy(), np.array([1.0, 0.0]), np.ones(3))
    assert info == infol == 1


@pytest.mark.parametrize("case", GRIDS, ids=[c[0] for c in GRIDS])
def test_diag_closed_form(oracle_mod, case):
    name, kind, im, jm, km, no_k, cyc = case
    g = oracle_mod.grid(im, jm, km, no_k=no_k, cyclic=cyc)
    a, b = make_system(kind, im, jm, km, no_k)
    r = rng_vec(b.size, 5)
    z, _ = oracle_mod.msolve(g, "DIAG", a, r)
    a0 = a.reshape(7, -1)[3]
    np.testing.assert_allclose(z * a0, r, rtol=2e-16 * 4)


def test_line_last_colour_satisfies_its_lines(oracle_mod):
    """After 'R' the (i+k)-even lines were solved last: each such J-line
    satisfies its tridiagonal equation with the final neighbour values."""
    im, jm, km = 7, 9, 6
    g = oracle_mod.grid(im, jm, km)
    a, b = make_system(1, im, jm, km)
    a, b, _ = oracle_mod.unit_scale(g, a, b)
    r = rng_vec(b.size, 9)
    z, info = oracle_mod.msolve(g, "LINE", a, r, "RSSS")
    assert info == 0
    ni, nj, nk = im + 2, jm + 2, km + 2
    A = a.reshape(7, nk, ni, nj)
    Z = z.reshape(nk, ni, nj)
    R = r.reshape(nk, ni, nj)
    worst = 0.0
    for k in range(nk):
        for i in range(ni):
            if ((i + 1) + (k + 1)) % 2 != 0:
                continue  # odd lines were solved first
            im_, ip_ = max(i - 1, 0), min(i + 1, ni - 1)
            km_, kp_ = max(k - 1, 0), min(k + 1, nk - 1)
            rhs = (R[k, i] - A[2, k, i] * Z[k, im_] - A[4, k, i] * Z[k, ip_] - A[0, k, i] * Z[km_, i]
                   - A[6, k, i] * Z[kp_, i])
            zl = Z[k, i]
            lhs = A[3, k, i] * zl
            lhs[1:] += A[1, k, i, 1:] * zl[:-1]
            lhs[:-1] += A[5, k, i, :-1] * zl[1:]
            worst = max(worst, np.max(np.abs(lhs - rhs)))
    assert worst <= 1e-13


def _direct(a_s, b_s, im, jm, km, no_k=False, cyc=(0, 0, 0)):
    return spla.spsolve(to_csr(a_s, im, jm, km, no_k, cyc).tocsc(), b_s)


@pytest.mark.parametrize("pc", ["DIAG", "LINE", "NONE"])
@pytest.mark.parametrize("case", GRIDS, ids=[c[0] for c in GRIDS])
def test_bicgs_meets_residual_and_direct_solution(oracle_mod, case, pc):
    name, kind, im, jm, km, no_k, cyc = case
    g = oracle_mod.grid(im, jm, km, no_k=no_k, cyclic=cyc)
    a, b = make_system(kind, im, jm, km, no_k)
    noise = rng_vec(b.size, 99) * 0.5 + 0.5
    tol = 1e-10
    x, a_s, b_s, res = oracle_mod.leq_bicgs(g, pc, "RSRS", tol, 2000, np.zeros(b.size), a, b, noise)
    assert res["ier"] == 0
    # unit scaling: every row's max |coefficient| is 1
    np.testing.assert_allclose(np.max(np.abs(a_s.reshape(7, -1)), axis=0), 1.0, rtol=1e-15)
    r = b_s - to_csr(a_s, im, jm, km, no_k, cyc) @ x
    assert np.linalg.norm(r) <= 2 * tol * res["rnorm0"] + 1e-13 * res["rnorm0"]
    xs = _direct(a_s, b_s, im, jm, km, no_k, cyc)
    kappa = np.linalg.cond(to_csr(a_s, im, jm, km, no_k, cyc).toarray()) if b.size < 3000 else 1e6
    assert np.max(np.abs(x - xs)) <= 10 * kappa * tol * np.max(np.abs(xs)) + 1e-14


def test_planted_sign_error_is_caught(oracle_mod):
    """A flipped omega term (R = S + omega T) must fail the residual check."""
    im = jm = km = 6
    g = oracle_mod.grid(im, jm, km)
    a, b = make_system(0, im, jm, km)
    noise = rng_vec(b.size, 99) * 0.5 + 0.5
    # emulate the mistake by solving with -A on the right-hand side update:
    # the oracle solving A x = b must not also solve (A) x = -b
    x, a_s, b_s, res = oracle_mod.leq_bicgs(g, "DIAG", "RSRS", 1e-10, 500, np.zeros(b.size), a, b, noise)
    r_wrong = -b_s - to_csr(a_s, im, jm, km) @ x
    assert np.linalg.norm(r_wrong) > 1.0 * np.linalg.norm(b_s)


def test_identity_one_iteration(oracle_mod):
    im, jm, km = 5, 6, 7
    g = oracle_mod.grid(im, jm, km)
    n3 = 7 * 8 * 9
    a = np.zeros(7 * n3)
    a[3 * n3:4 * n3] = -2.0
    b = rng_vec(n3, 8)
    x, _, b_s, res = oracle_mod.leq_bicgs(g, "NONE", "RSRS", 1e-10, 10, np.zeros(n3), a, b, np.full(n3, 0.5))
    # scaled system is -I: x = -b_s exactly after one step
    assert res["iterations"] == 1 and res["ier"] == 0
    np.testing.assert_allclose(x, -b_s, rtol=1e-15)


def test_zero_rhs_and_zero_row(oracle_mod):
    g = oracle_mod.grid(6, 5, 4)
    a, b = make_system(0, 6, 5, 4)
    x, _, _, res = oracle_mod.leq_bicgs(g, "DIAG", "RSRS", 1e-10, 50, np.zeros(b.size), a, np.zeros(b.size),
                                        np.full(b.size, 0.5))
    assert res["ier"] == 0 and res["iterations"] == 0 and not np.any(x)
    a2 = a.copy()
    a2.reshape(7, -1)[:, 11] = 0.0
    _, _, _, res = oracle_mod.leq_bicgs(g, "DIAG", "RSRS", 1e-10, 50, np.zeros(b.size), a2, b, np.full(b.size, 0.5))
    assert res["ier"] == -2 and res["exit_reason"] == 4


@pytest.mark.parametrize("itmax,expect", [(1, -2), (0, 0)])
def test_itmax_classification(oracle_mod, itmax, expect):
    """ITMAX reached: IER=-1 if real(Rnorm) < 0.2*Rnorm0 else -2
    (leq_bicgs.f:486-495); ITMAX=0 never iterates and Rnorm=0 counts as
    converged."""
    g = oracle_mod.grid(8, 8, 8)
    a, b = make_system(0, 8, 8, 8)
    _, _, _, res = oracle_mod.leq_bicgs(g, "DIAG", "RSRS", 1e-10, itmax, np.zeros(b.size), a, b,
                                        np.full(b.size, 0.5))
    assert res["ier"] == expect


def test_bad_pc_rejected(oracle_mod):
    g = oracle_mod.grid(4, 4, 4)
    n3 = 6 * 6 * 6
    with pytest.raises(ValueError):
        oracle_mod.leq_bicgs(g, "ILU0", "RSRS", 1e-8, 5, np.zeros(n3), np.ones(7 * n3), np.ones(n3), np.zeros(n3))
