"""GPU CALC_RESID_C / _S / _PP (model/calc_resid.f) against the oracle,
through the C ABI (device and host entries).  Per-cell residuals and the
location of the largest one are computed as the reference does, so
IJK_RESID and the fluid-cell count match exactly; NUM and DEN are sums
taken in another order than the reference's serial loop: relative 1e-13."""
import numpy as np
import pytest

import oracle
from tests.resid_cases import flags_for, resid_cases

pytestmark = pytest.mark.gpu

CASES = resid_cases()


def _check(got, want):
    assert got["ijk_resid"] == want["ijk_resid"], (got, want)
    assert got["ncells"] == want["ncells"], (got, want)
    for k in ("num", "den", "resid", "max_resid"):
        np.testing.assert_allclose(got[k], want[k], rtol=1e-13, atol=0.0, err_msg=k)


def _oracle(case):
    return oracle.calc_resid(case["grid"], case["kind"], case["x"], case["a"], case["b"], case["flag"],
                             tol=case["tol"], ep_s=case["ep_s"] if case["m"] > 0 else None, norm=case["norm"],
                             debug=case["debug"])


def _grid_init(gpu, g):
    gpu.grid_init(g.imax, g.jmax, g.kmax, no_k=bool(g.no_k),
                  cyclic=(bool(g.cyclic_x), bool(g.cyclic_y), bool(g.cyclic_z)))


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_resid_device_matches_oracle(gpu, case):
    _grid_init(gpu, case["grid"])
    dev = {k: gpu.DeviceArray.from_host(np.ascontiguousarray(case[k])) for k in ("x", "a", "b")}
    dflag = gpu.DeviceArray.from_host(np.ascontiguousarray(case["flag"], dtype=np.int32))
    deps = gpu.DeviceArray.from_host(case["ep_s"]) if case["m"] > 0 else None
    got = gpu.calc_resid_dev(case["kind"], b=dev["b"], flag=dflag, var=dev["x"], a=dev["a"], tol=case["tol"],
                             ep_s=deps, norm=case["norm"], debug=case["debug"])
    _check(got, _oracle(case))


@pytest.mark.parametrize("kind", ["C", "S", "PP"])
def test_resid_host_entry_larger_grid(gpu, kind):
    im, jm, km = 40, 36, 30
    from tests.cases import make_system
    a, b = make_system(0, im, jm, km, False, seed=31)
    n3 = b.size
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, n3)
    flag = flags_for(im, jm, km, False, 77)
    ep_s = rng.uniform(0, 2e-4, n3) if kind == "S" else None
    g = oracle.grid(im, jm, km)
    _grid_init(gpu, g)
    for debug in (True, False):
        got = gpu.calc_resid(kind, b=b, flag=flag, var=x, a=a, tol=0.2, ep_s=ep_s, norm=0.5, debug=debug)
        want = oracle.calc_resid(g, kind, x, a, b, flag, tol=0.2, ep_s=ep_s, norm=0.5, debug=debug)
        _check(got, want)
