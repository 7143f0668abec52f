"""LBQN quasi-Newton mode (src/quasi_newton.jl:15-58, src/algorithms_outer.jl:132-136)
through the C ABI against the oracle: the history shift and new columns
bit-exact, the step (Gram-form recursion on the device, the reference's
sequential two-loop in the oracle) to 1e-9, and whole admixture_qn!(mode=:LBQN)
iterations to the LL bar of 1e-10."""
import numpy as np
import pytest

from conftest import make_cohort

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("it,Q", [(1, 3), (2, 3), (5, 3), (7, 5)])
def test_lbqn_step(oa, O, it, Q):
    print("lbqn step", it, Q, flush=True)
    I, J, K = 300, 700, 4
    rng = np.random.default_rng(it * 10 + Q)
    x = np.asfortranarray(rng.random((K, I + J)))
    xn = np.asfortranarray(x + 0.01 * rng.normal(size=x.shape))
    xn2 = np.asfortranarray(xn + 0.008 * rng.normal(size=x.shape))
    U0 = np.asfortranarray(0.01 * rng.normal(size=(K * (I + J), Q)))
    V0 = np.asfortranarray(0.01 * rng.normal(size=(K * (I + J), Q)))
    c = oa.HIPAdmixData(I, J, K, Q)
    gU, gV = U0.copy(order="F"), V0.copy(order="F")
    gx = np.zeros_like(x, order="F")
    c.lbqn_step(it, x, xn, xn2, gU, gV, gx)
    rU, rV = U0.copy(order="F"), V0.copy(order="F")
    O.update_UV_LBQN(rU, rV, x, xn, xn2, it, Q)
    rx = np.zeros_like(x, order="F")
    O.update_QN_LBQN(rx, x, rU, rV, min(it, Q))
    O.project_p(rx[:, I:])
    rxq = np.asfortranarray(rx[:, :I])
    O.project_q(rxq)
    rx[:, :I] = rxq
    assert np.array_equal(gU, rU) and np.array_equal(gV, rV)
    assert np.abs(gx - rx).max() < 1e-9, np.abs(gx - rx).max()
    c.close()


def test_lbqn_iterations_match_oracle(oa, O):
    print("lbqn iterations", flush=True)
    I, J, K = 700, 2600, 4
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, seed=21, miss=0.01)
    d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(121))
    c = oa.HIPAdmixData(I, J, K, 3)
    c.upload_bed(bed.reshape((pitch, J), order="F"))
    ref = O.OracleAdmix(bed, pitch, I, J, K, 3, d.x, d.U, d.V)
    oa.init_em(d, None, 3, d_cu=c, log=None)
    ref.init_em(3)
    c.set_qn_mode("LBQN")
    for it in range(1, 9):
        r = c.qn_iteration(it)
        rr = ref.qn_iteration(it, mode="LBQN")
        for key in ("basic", "qn", "new"):
            assert abs(r[key] - rr[key]) <= 1e-10 * abs(rr[key]), (it, key, r[key], rr[key])
    assert np.abs(c.get_x(0) - ref.x).max() < 1e-6
    c.close()


def test_admixture_qn_lbqn_converges_to_the_zal_optimum(oa, O):
    print("lbqn convergence", flush=True)
    I, J, K = 900, 4000, 3
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, seed=5, miss=0.01)
    res = {}
    for mode in ("ZAL", "LBQN"):
        c = oa.HIPAdmixData(I, J, K, 3)
        c.upload_bed(bed.reshape((pitch, J), order="F"))
        d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(7))
        oa.init_em(d, None, 5, d_cu=c, log=None)
        hist = oa.admixture_qn(d, None, 500, 1e-9, d_cu=c, mode=mode, log=None)
        res[mode] = (hist[-1], len(hist) - 1)
        c.close()
    print(res, flush=True)
    assert abs(res["LBQN"][0] - res["ZAL"][0]) <= 1e-7 * abs(res["ZAL"][0])
