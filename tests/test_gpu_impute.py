"""Mean-imputation loop family (skipmissing = false; src/loops.jl:493-514,
589-608, 679-698, 840-865, 896-921) through the C ABI against the oracle: the
five tile loops over ragged ranges, and init_em!/admixture_qn! on
AdmixData(skipmissing=false), to the same bars as the skipmissing family."""
import numpy as np
import pytest

from conftest import make_cohort, random_state

pytestmark = pytest.mark.gpu


def _colrel(got, ref):
    got = np.asarray(got).reshape(-1, np.asarray(got).shape[-1])
    ref = np.asarray(ref).reshape(-1, np.asarray(ref).shape[-1])
    return (np.abs(got - ref).max(0) / (np.abs(ref).max(0) + 1e-300)).max()


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_imputation_loops_match_oracle(oa, O, T):
    print("impute loops", np.dtype(T).name, flush=True)
    I, J, K = 700, 1300, 4
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.08, seed=41)
    c = oa.HIPAdmixData(I, J, K, 3, T)
    c.upload_bed(bed.reshape((pitch, J), order="F"))
    g = c
    q, p = random_state(I, J, K, 7, T)
    mu = O.snp_means(bed, pitch, I, J, dtype=T)
    ll_tol = 1e-10 if T == np.float64 else 1e-5
    h_tol = 1e-9 if T == np.float64 else 2e-4
    for (i0, i1, j0, j1) in [(0, I, 0, J), (3, I - 5, 7, J - 2)]:
        ir, jr = oa.UnitRange(i0 + 1, i1), oa.UnitRange(j0 + 1, j1)
        got = oa.loglikelihood_loop(g, q, p, None, ir, jr, K)
        ref = O.loglikelihood(bed, pitch, q, p, (i0, i1), (j0, j1), mu=mu)
        assert abs(got - ref) <= ll_tol * abs(ref), (got, ref)
        gX, gZ = np.zeros((K, K, I), T, order="F"), np.zeros((K, I), T, order="F")
        rX, rZ = np.zeros((K, K, I), T, order="F"), np.zeros((K, I), T, order="F")
        oa.update_q_loop(gX, gZ, g, q, p, None, ir, jr, K)
        O.update_q_loop(rX, rZ, bed, pitch, q, p, (i0, i1), (j0, j1), mu=mu)
        assert _colrel(gX[:, :, i0:i1], rX[:, :, i0:i1]) < h_tol and _colrel(gZ[:, i0:i1], rZ[:, i0:i1]) < h_tol
        gX, gZ = np.zeros((K, K, J), T, order="F"), np.zeros((K, J), T, order="F")
        rX, rZ = np.zeros((K, K, J), T, order="F"), np.zeros((K, J), T, order="F")
        oa.update_p_loop(gX, gZ, g, q, p, None, ir, jr, K)
        O.update_p_loop(rX, rZ, bed, pitch, q, p, (i0, i1), (j0, j1), mu=mu)
        assert _colrel(gX[:, :, j0:j1], rX[:, :, j0:j1]) < h_tol and _colrel(gZ[:, j0:j1], rZ[:, j0:j1]) < h_tol
        gq, rq = np.zeros((K, I), T, order="F"), np.zeros((K, I), T, order="F")
        oa.em_q_loop(gq, g, q, p, None, ir, jr, K)
        O.em_q_loop(rq, bed, pitch, q, p, (i0, i1), (j0, j1), mu=mu)
        assert _colrel(gq[:, i0:i1], rq[:, i0:i1]) < h_tol
        gpn, gpt = np.zeros((K, J), T, order="F"), np.zeros((K, J), T, order="F")
        rpn, rpt = np.zeros((K, J), T, order="F"), np.zeros((K, J), T, order="F")
        oa.em_p_loop(gpn, g, q, p, gpt, None, ir, jr, K)
        O.em_p_loop(rpn, rpt, bed, pitch, q, p, (i0, i1), (j0, j1), mu=mu)
        assert _colrel(gpn[:, j0:j1], rpn[:, j0:j1]) < h_tol and _colrel(gpt[:, j0:j1], rpt[:, j0:j1]) < h_tol
    # the context is back on the skipmissing family afterwards
    assert abs(c.loglikelihood_loop(q, p, 0, I, 0, J) - O.loglikelihood(bed, pitch, q, p)) <= ll_tol * abs(
        O.loglikelihood(bed, pitch, q, p))
    c.close()


def test_imputation_iterations_match_oracle(oa, O):
    print("impute iterations", flush=True)
    I, J, K = 600, 2200, 3
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.06, seed=42)
    d = oa.AdmixData(I, J, K, 3, skipmissing=False, rng=np.random.default_rng(5))
    c = oa.HIPAdmixData(I, J, K, 3)
    c.upload_bed(bed.reshape((pitch, J), order="F"))
    ref = O.OracleAdmix(bed, pitch, I, J, K, 3, d.x, d.U, d.V, mu=O.snp_means(bed, pitch, I, J))
    lls = oa.init_em(d, None, 3, d_cu=c, log=None)
    np.testing.assert_allclose(lls, ref.init_em(3), rtol=1e-10, atol=0)
    for it in range(1, 7):
        r = c.qn_iteration(it)
        rr = ref.qn_iteration(it)
        for key in ("basic", "qn", "new"):
            assert abs(r[key] - rr[key]) <= 1e-10 * abs(rr[key]), (it, key, r[key], rr[key])
    assert np.abs(c.get_x(0) - ref.x).max() < 1e-6
    c.set_precision("bf16")
    with pytest.raises(oa.AdmixHIPError):
        c.update_q(False)
    c.close()
