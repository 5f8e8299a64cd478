"""The algebra of the FP32 several-lanes-per-line solve (CPU only).

k_line_pcr_sys splits a line of n rows over P lanes of m = ceil(n/P) rows:
each lane eliminates its rows but the last, expressing them through its own
last row (the separator) and the previous lane's; the separators form a
tridiagonal system solved by parallel cyclic reduction; each lane then
recovers its rows (partition method, Wang 1981; PCR, Hockney 1965).  This
test restates those steps in numpy, lane by lane as the kernel orders them,
and checks them in FP64 against scipy's banded LU on diagonally dominant
lines of every length class the kernel's variants take, including lines
shorter than P (idle lanes) and lengths that fill the lanes exactly.  It
also checks the kernel's LDS layout: rows are loaded in interleaved order
(row q by lane q mod P) and read back by chunk (rows gl*m .. gl*m+m-1 by
lane gl), through slots l*s + (q - l*m) with l = q // m and s = m | 1.
"""
import numpy as np
import pytest
from scipy.linalg import solve_banded

VARIANTS = [(16, 2), (16, 4), (32, 4), (32, 5), (64, 4), (64, 5), (64, 8)]


def partition_pcr(a, b, c, d, P, M):
    n = len(b)
    m = (n + P - 1) // P
    assert m <= M
    # rows past n: identity rows (x = 0, no coupling)
    pad = P * m
    A = np.zeros(pad); B = np.ones(pad); C = np.zeros(pad); D = np.zeros(pad)
    A[:n], B[:n], C[:n], D[:n] = a, b, c, d
    A, B, C, D = (v.reshape(P, m) for v in (A, B, C, D))
    cp, wp, dp = np.zeros((P, m)), np.zeros((P, m)), np.zeros((P, m))
    U, V, W = np.zeros((P, m)), np.zeros((P, m)), np.zeros((P, m))
    for j in range(m - 1):
        if j == 0:
            r = 1.0 / B[:, 0]
            cp[:, 0], wp[:, 0], dp[:, 0] = C[:, 0] * r, A[:, 0] * r, D[:, 0] * r
        else:
            r = 1.0 / (B[:, j] - A[:, j] * cp[:, j - 1])
            cp[:, j] = C[:, j] * r
            wp[:, j] = -A[:, j] * wp[:, j - 1] * r
            dp[:, j] = (D[:, j] - A[:, j] * dp[:, j - 1]) * r
    for j in range(m - 2, -1, -1):
        if j == m - 2:
            U[:, j], V[:, j], W[:, j] = dp[:, j], cp[:, j], wp[:, j]
        else:
            U[:, j] = dp[:, j] - cp[:, j] * U[:, j + 1]
            V[:, j] = -cp[:, j] * V[:, j + 1]
            W[:, j] = wp[:, j] - cp[:, j] * W[:, j + 1]
    a_s, b_s, c_s, d_s = A[:, m - 1], B[:, m - 1], C[:, m - 1], D[:, m - 1]
    if m > 1:
        Un, Vn, Wn = (np.append(X[1:, 0], 0.0) for X in (U, V, W))
        RA = -a_s * W[:, m - 2]
        RB = b_s - a_s * V[:, m - 2] - c_s * Wn
        RC = -c_s * Vn
        RD = d_s - a_s * U[:, m - 2] - c_s * Un
    else:
        RA, RB, RC, RD = a_s.copy(), b_s.copy(), c_s.copy(), d_s.copy()
    lane = np.arange(P)
    s = 1
    while s < P:
        hl, hr = lane >= s, lane + s < P
        il, ir = np.where(hl, lane - s, lane), np.where(hr, lane + s, lane)
        al = np.where(hl, -RA / RB[il], 0.0)
        ga = np.where(hr, -RC / RB[ir], 0.0)
        RA, RB, RC, RD = (np.where(hl, al * RA[il], 0.0),
                          RB + np.where(hl, al * RC[il], 0.0) + np.where(hr, ga * RA[ir], 0.0),
                          np.where(hr, ga * RC[ir], 0.0),
                          RD + np.where(hl, al * RD[il], 0.0) + np.where(hr, ga * RD[ir], 0.0))
        s *= 2
    G = RD / RB
    Gp = np.append(0.0, G[:-1])
    x = np.empty((P, m))
    for j in range(m):
        x[:, j] = U[:, j] - V[:, j] * G - W[:, j] * Gp if j < m - 1 else G
    return x.reshape(-1)[:n]


@pytest.mark.parametrize("P,M", VARIANTS)
def test_partition_pcr_solves_the_line(P, M):
    rng = np.random.default_rng(P * 10 + M)
    for n in sorted({1, 2, 3, P - 1, P, P + 1, P * M // 2 + 3, P * M - 1, P * M}):
        if n < 1 or (n + P - 1) // P > M:
            continue
        a = rng.uniform(0.1, 1.0, n)
        c = rng.uniform(0.1, 1.0, n)
        a[0], c[-1] = 0.0, 0.0
        b = -(a + c) - rng.uniform(0.01, 0.5, n)
        d = rng.uniform(-1.0, 1.0, n)
        ab = np.zeros((3, n))
        ab[0, 1:], ab[1], ab[2, :-1] = c[:-1], b, a[1:]
        x_ref = solve_banded((1, 1), ab, d)
        x = partition_pcr(a, b, c, d, P, M)
        assert np.max(np.abs(x - x_ref)) <= 1e-12 * np.max(np.abs(x_ref)), (P, M, n)


def test_planted_sign_error_is_caught():
    """A sign slip in the separator equation (the kind of mistake the
    restatement must expose) breaks the agreement."""
    rng = np.random.default_rng(5)
    n = 100
    a, c = rng.uniform(0.1, 1, n), rng.uniform(0.1, 1, n)
    a[0] = c[-1] = 0.0
    b = -(a + c) - 0.1
    d = rng.uniform(-1, 1, n)
    x = partition_pcr(a, b, c, d, 16, 8)
    y = b * x + np.append(0.0, a[1:] * x[:-1]) + np.append(c[:-1] * x[1:], 0.0)
    assert np.max(np.abs(y - d)) < 1e-12
    x_bad = partition_pcr(a, b, -c, d, 16, 8)
    y_bad = b * x_bad + np.append(0.0, a[1:] * x_bad[:-1]) + np.append(c[:-1] * x_bad[1:], 0.0)
    assert np.max(np.abs(y_bad - d)) > 1e-3


def _slot_f32(q, m):
    """The kernel's slot of row q: q // m from one FP32 multiply by 1/m."""
    rm = np.float32(1.0) / np.float32(m)
    l = int(np.float32(np.float32(q) + np.float32(0.5)) * rm)
    return l * (m | 1) + (q - l * m), l


@pytest.mark.parametrize("P,M", VARIANTS)
def test_lds_slots_transpose_rows_to_chunks(P, M):
    """Every row of every length the variant takes lands in its own slot
    below P*(M+1), a lane's chunk reads find exactly its rows, and the
    chunk reads of the P lanes fall in P distinct banks of 64."""
    for n in range(1, P * M + 1):
        m = (n + P - 1) // P
        if m > M:
            continue
        slots = {}
        for q in range(n):
            o, l = _slot_f32(q, m)
            assert l == q // m, (P, M, n, q)
            assert 0 <= o < P * (M + 1)
            slots[o] = q
        assert len(slots) == n
        for gl in range(P):
            for j in range(m):
                q = gl * m + j
                if q < n:
                    assert slots[gl * (m | 1) + j] == q
        for j in range(m):
            banks = {(gl * (m | 1) + j) % 64 for gl in range(P)}
            assert len(banks) == P


def test_slot_division_exact_for_every_row_index():
    """(q + 0.5) * fl(1/m) in FP32 truncates to q // m for q < 4096, m <= 8
    (the kernel's lines have at most 512 rows)."""
    for m in range(1, 9):
        q = np.arange(4096, dtype=np.float32)
        l = ((q + np.float32(0.5)) * (np.float32(1.0) / np.float32(m))).astype(np.int64)
        assert np.array_equal(l, np.arange(4096) // m), m
