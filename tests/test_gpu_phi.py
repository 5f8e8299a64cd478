"""GPU scalar-transport assembly (INIT_AB_M + CONV_DIF_PHI with first-order
upwinding, SOURCE_PHI, UNDER_RELAX_S; model/conv_dif_phi.f,
model/source_phi.f, model/under_relax.f) against the oracle and the
reference's golden hashes, through the C ABI.  All three are element-wise
arithmetic in the reference's order: A_M and B_M must match bit for bit."""
import hashlib
from pathlib import Path

import numpy as np
import pytest

import oracle
from tests.phi_cases import DIL_EP_S, phi_cases, phi_fields, phi_flags, reference_chain

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "phi_ref_golden.npz"
CASES = phi_cases()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _grid_init(gpu, g):
    gpu.grid_init(g.imax, g.jmax, g.kmax, no_k=bool(g.no_k),
                  cyclic=(bool(g.cyclic_x), bool(g.cyclic_y), bool(g.cyclic_z)))


def _device_run(gpu, case):
    fields = {k: gpu.DeviceArray.from_host(np.ascontiguousarray(v)) for k, v in case["fields"].items()}
    flag = gpu.DeviceArray.from_host(np.ascontiguousarray(case["flag"], dtype=np.int32))
    n3 = oracle.n3_of(case["grid"])

    def run(stage, a, b):
        if stage == 1:
            # stale contents: the kernel must write every entry
            da = gpu.DeviceArray.from_host(np.full(7 * n3, 5.0))
            db = gpu.DeviceArray.from_host(np.full(n3, 5.0))
            gpu.conv_dif_phi_dev(fields, flag, da, db)
        else:
            da, db = gpu.DeviceArray.from_host(a), gpu.DeviceArray.from_host(b)
            if stage == 2:
                gpu.source_phi_dev(fields["s_p"], fields["s_c"], fields["ep"], fields["phi"], flag, da, db,
                                   dil_ep_s=DIL_EP_S)
            else:
                gpu.under_relax_s_dev(fields["phi"], flag, case["ur"], da, db)
        return da.download(), db.download()
    return run


def _oracle_run(case):
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
def test_chain_bit_exact(gpu, golden, case):
    _grid_init(gpu, case["grid"])
    got = reference_chain(case, _device_run(gpu, case))
    want = reference_chain(case, _oracle_run(case))
    for s, ((a, b), (ao, bo)) in enumerate(zip(got, want), start=1):
        np.testing.assert_array_equal(a, ao, err_msg=f"A_M after stage {s}")
        np.testing.assert_array_equal(b, bo, err_msg=f"B_M after stage {s}")
        assert sha(a) == str(golden[f"{case['name']}/a{s}_sha"])
        assert sha(b) == str(golden[f"{case['name']}/b{s}_sha"])


@pytest.mark.parametrize("dims,cyc", [((64, 48, 40), (0, 0, 0)), ((50, 41, 37), (1, 0, 1)), ((90, 70, 1), (0, 0, 0))])
def test_chain_bit_exact_larger(gpu, dims, cyc):
    im, jm, km = dims
    no_k = km == 1
    g = oracle.grid(im, jm, km, no_k=no_k, cyclic=cyc)
    n3 = oracle.n3_of(g)
    ni, nj, nk = im + 2, jm + 2, 1 if no_k else km + 2
    fields, b_in = phi_fields(n3, ni, nj, nk, no_k, 77)
    case = dict(grid=g, flag=phi_flags(im, jm, km, no_k, cyc, 76), fields=fields, b_in=b_in, ur=0.7)
    _grid_init(gpu, g)
    got = reference_chain(case, _device_run(gpu, case))
    want = reference_chain(case, _oracle_run(case))
    for s, ((a, b), (ao, bo)) in enumerate(zip(got, want), start=1):
        np.testing.assert_array_equal(a, ao, err_msg=f"A_M after stage {s}")
        np.testing.assert_array_equal(b, bo, err_msg=f"B_M after stage {s}")


def test_under_relax_rejects_zero_factor(gpu):
    case = CASES[0]
    g = case["grid"]
    n3 = oracle.n3_of(g)
    _grid_init(gpu, g)
    da, db = gpu.DeviceArray(7 * n3), gpu.DeviceArray(n3)
    var = gpu.DeviceArray(n3)
    flag = gpu.DeviceArray.from_host(np.ascontiguousarray(case["flag"], dtype=np.int32))
    from mfix_amd._lib import LeqError
    with pytest.raises(LeqError):
        gpu.under_relax_s_dev(var, flag, 0.0, da, db)
