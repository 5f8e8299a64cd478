"""The oracle's scalar-transport assembly (INIT_AB_M + CONV_DIF_PHI with
first-order upwinding, SOURCE_PHI, UNDER_RELAX_S; model/conv_dif_phi.f,
model/source_phi.f, model/under_relax.f) against the compiled reference's
outputs, bit for bit (tests/golden/phi_ref_golden.npz, written by
tests/golden/make_phi_golden.py), and against what the discretisation fixes:
the upwind split of each face flux, the harmonic mean's closed forms, the
source rows' balance and the under-relaxation identities."""
import hashlib
from pathlib import Path

import numpy as np
import pytest

import oracle
from tests.phi_cases import DIL_EP_S, phi_cases, phi_input_sha, reference_chain

GOLDEN = Path(__file__).resolve().parent / "golden" / "phi_ref_golden.npz"
CASES = phi_cases()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def oracle_run(case):
    g, f, fl = case["grid"], case["fields"], case["flag"]

    def run(stage, a, b):
        if stage == 1:
            return oracle.conv_dif_phi(g, f, fl)
        if stage == 2:
            return oracle.source_phi(g, a, b, f["s_p"], f["s_c"], f["ep"], f["phi"], fl, DIL_EP_S)
        return oracle.under_relax_s(g, a, b, f["phi"], fl, case["ur"])
    return run


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_chain_matches_reference_golden(golden, case):
    key = case["name"]
    assert str(golden[f"{key}/in_sha"]) == phi_input_sha(case)
    for s, (a, b) in enumerate(reference_chain(case, oracle_run(case)), start=1):
        assert sha(a) == str(golden[f"{key}/a{s}_sha"]), f"A_M after stage {s}"
        assert sha(b) == str(golden[f"{key}/b{s}_sha"]), f"B_M after stage {s}"


def _interior_pairs(case, axis):
    """(c, e) of fluid cells and their plain +1 neighbour along axis, both fluid."""
    g = case["grid"]
    ni, nj = g.imax + 2, g.jmax + 2
    nk = 1 if g.no_k else g.kmax + 2
    fl = (case["flag"] == 1).reshape(nk, ni, nj)
    idx = np.arange(ni * nj * nk).reshape(nk, ni, nj)
    if axis == 0:
        m = fl[:, :-1, :] & fl[:, 1:, :]
        return idx[:, :-1, :][m], idx[:, 1:, :][m]
    if axis == 1:
        m = fl[:, :, :-1] & fl[:, :, 1:]
        return idx[:, :, :-1][m], idx[:, :, 1:][m]
    m = fl[:-1] & fl[1:]
    return idx[:-1][m], idx[1:][m]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_upwind_split_of_each_face(case):
    """Between two fluid cells the east entry of one and the west entry of the
    other are D and D + F (or D - F and D): their difference is the face flux,
    the upwind side takes it, and both are >= D's share of zero."""
    g, f = case["grid"], case["fields"]
    n3 = oracle.n3_of(g)
    A = oracle.conv_dif_phi(g, f, case["flag"])[0].reshape(7, n3)
    axes = [(0, 4, 2, "u", "flux_e"), (1, 5, 1, "v", "flux_n")]
    if not g.no_k:
        axes.append((2, 6, 0, "w", "flux_t"))
    for axis, plus, minus, vel, flux in axes:
        c, e = _interior_pairs(case, axis)
        assert c.size > 0
        F = f[flux][c]
        diff = A[minus, e] - A[plus, c]
        np.testing.assert_allclose(diff, F, rtol=1e-12, atol=1e-12 * np.abs(A[plus, c]).max())
        up = f[vel][c] >= 0
        # the downstream side's entry is the bare diffusion term
        d_down = np.where(up, A[plus, c], A[minus, e])
        np.testing.assert_allclose(np.where(up, A[minus, e] - F, A[plus, c] + F), d_down, rtol=1e-12,
                                   atol=1e-20)


def test_harmonic_mean_closed_forms():
    """Uniform DIF and central weights: D = DIF * ODX_E * AYZ; DIF = 0 on one
    side: D = 0 (the SMALL_NUMBER floor keeps the division finite)."""
    case = dict(CASES[0])
    g = case["grid"]
    n3 = oracle.n3_of(g)
    f = {k: v.copy() for k, v in case["fields"].items()}
    f["dif"][:] = 1.5e-3
    f["fx_e"][:] = 0.5
    f["fx_e_bar"][:] = 0.5
    f["u"][:] = 1.0           # every east entry is the bare diffusion term
    A = oracle.conv_dif_phi(g, f, case["flag"])[0].reshape(7, n3)
    c, e = _interior_pairs(case, 0)
    ni_nj = (g.imax + 2) * (g.jmax + 2)
    i = (c % ni_nj) // (g.jmax + 2)
    np.testing.assert_allclose(A[4, c], 1.5e-3 * f["odx_e"][i] * f["ayz"][c], rtol=1e-15)
    f["dif"][e] = 0.0
    A = oracle.conv_dif_phi(g, f, case["flag"])[0].reshape(7, n3)
    assert np.isfinite(A).all()
    assert (A[4, c] == 0.0).all()


@pytest.mark.parametrize("case", CASES[:5], ids=[c["name"] for c in CASES[:5]])
def test_source_rows(case):
    """SOURCE_PHI: a dense fluid row gets A0 = -(sum of neighbours + S_P),
    less B/PHI when B >= S_C and PHI != 0, and B = B - S_C or -S_C; a dilute
    row couples with weight 1 to its non-dilute neighbours only; non-fluid
    rows are untouched."""
    g, f, fl = case["grid"], case["fields"], case["flag"] == 1
    n3 = oracle.n3_of(g)
    a1, _ = oracle.conv_dif_phi(g, f, case["flag"])
    b_in = case["b_in"]
    a2, b2 = oracle.source_phi(g, a1, b_in, f["s_p"], f["s_c"], f["ep"], f["phi"], case["flag"], DIL_EP_S)
    A1, A2 = a1.reshape(7, n3), a2.reshape(7, n3)
    off = [0, 1, 2, 4, 5, 6]
    np.testing.assert_array_equal(A2[:, ~fl], A1[:, ~fl])
    np.testing.assert_array_equal(b2[~fl], b_in[~fl])
    dense = fl & (f["ep"] > DIL_EP_S)
    branch1 = dense & ((b_in < f["s_c"]) | (f["phi"] == 0))
    branch2 = dense & ~branch1
    total = A1[off].sum(axis=0) + f["s_p"]
    np.testing.assert_allclose(A2[3, branch1], -total[branch1], rtol=1e-13)
    np.testing.assert_allclose(b2[branch1], b_in[branch1] - f["s_c"][branch1], rtol=1e-13, atol=1e-20)
    np.testing.assert_allclose(A2[3, branch2], -total[branch2] - b_in[branch2] / f["phi"][branch2], rtol=1e-12)
    np.testing.assert_array_equal(b2[branch2], -f["s_c"][branch2])
    assert branch1.any() and branch2.any()
    dil = fl & ~dense
    assert dil.any()
    assert set(np.unique(A2[off][:, dil])) <= {0.0, 1.0}
    nsum = A2[off][:, dil].sum(axis=0)
    np.testing.assert_array_equal(A2[3, dil], np.where(nsum == 0, -1.0, -nsum))
    np.testing.assert_array_equal(b2[dil], np.where(nsum == 0, -f["phi"][dil], 0.0))


@pytest.mark.parametrize("case", CASES[:3], ids=[c["name"] for c in CASES[:3]])
def test_under_relaxation_identities(case):
    """UR = 1 leaves the system as it is; UR = 0.5 doubles A0 and adds A0*VAR
    to B, except on rows whose A0 is exactly -1 and on non-fluid rows."""
    g, f, fl = case["grid"], case["fields"], case["flag"] == 1
    n3 = oracle.n3_of(g)
    a1, _ = oracle.conv_dif_phi(g, f, case["flag"])
    a2, b2 = oracle.source_phi(g, a1, case["b_in"], f["s_p"], f["s_c"], f["ep"], f["phi"], case["flag"])
    a3, b3 = oracle.under_relax_s(g, a2, b2, f["phi"], case["flag"], 1.0)
    np.testing.assert_array_equal(a3, a2)
    np.testing.assert_array_equal(b3, b2)
    a3, b3 = oracle.under_relax_s(g, a2, b2, f["phi"], case["flag"], 0.5)
    A2, A3 = a2.reshape(7, n3), a3.reshape(7, n3)
    keep = ~fl | (A2[3] == -1.0)
    assert keep.any() and (~keep).any()
    np.testing.assert_array_equal(A3[:, keep], A2[:, keep])
    np.testing.assert_array_equal(b3[keep], b2[keep])
    ch = ~keep
    np.testing.assert_array_equal(A3[3, ch], 2.0 * A2[3, ch])
    np.testing.assert_array_equal(b3[ch], b2[ch] + A2[3, ch] * f["phi"][ch])
    np.testing.assert_array_equal(np.delete(A3, 3, axis=0), np.delete(A2, 3, axis=0))
