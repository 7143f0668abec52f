"""Mixed-precision forms of the device-resident update passes (FP64 context,
FP32 or BF16-MFMA passes, FP64 QP / log-likelihood / convergence test).

Per-output error bars against the plain FP64 path on the same state, written
per test:
  Hessian (XtX)   FP32: 1e-5, BF16: 5e-3   relative to the column maximum
  gradient (Xtz)  FP32: 1e-5, BF16: 1e-5   relative to the column maximum
  converged LL    |LL_form - LL_f64| / |LL_f64| < 1e-6 (the benchmark bar is 1e-4)
  BF16 EM step    q_next 2e-5 relative to the column maximum, p_next 2e-5 absolute
  BF16 LL         1e-6 relative (matrix-core s and t)
"""
import numpy as np
import pytest

from conftest import make_cohort

pytestmark = pytest.mark.gpu

TOL = {"f32": (1e-5, 1e-5), "bf16": (5e-3, 1e-5)}


def _colrel(got, ref):
    got = np.asarray(got, np.float64).reshape(-1, got.shape[-1])
    ref = np.asarray(ref, np.float64).reshape(-1, ref.shape[-1])
    return (np.abs(got - ref).max(0) / (np.abs(ref).max(0) + 1e-300)).max()


def _ctx(oa, O, I, J, K, seed, beta=True):
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, alpha=0.3, miss=0.02, seed=seed, p_lo=0.01, p_hi=0.99, beta=beta)
    c = oa.HIPAdmixData(I, J, K, 3)
    c.upload_bed(bed.reshape((pitch, J), order="F"))
    d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(seed))
    return c, d


@pytest.mark.parametrize("K", [2, 3, 8, 12])
@pytest.mark.parametrize("form", ["f32", "bf16"])
def test_mixed_hessians_match_f64(oa, O, K, form):
    print("mixed hessians", K, form, flush=True)
    I, J = 1500, 4100
    c, d = _ctx(oa, O, I, J, K, seed=K)
    oa.init_em(d, None, 3, d_cu=c, log=None)
    out = {}
    for mode in ("f64", form):
        c.set_precision(mode)
        c.update_q(False)
        XtXq, Xtzq = c.hessians(0)
        c.set_x(1, d.x)  # same q_next (= q) for both modes' update_p
        c.update_p(False)
        XtXp, Xtzp = c.hessians(1)
        out[mode] = (XtXq, Xtzq, XtXp, Xtzp, c.get_x(1))
    htol, gtol = TOL[form]
    f, m = out["f64"], out[form]
    eq, gq, ep, gp = _colrel(m[0], f[0]), _colrel(m[1], f[1]), _colrel(m[2], f[2]), _colrel(m[3], f[3])
    print(f"K={K} {form}: XtX_q {eq:.2e} Xtz_q {gq:.2e} XtX_p {ep:.2e} Xtz_p {gp:.2e}", flush=True)
    assert eq < htol and ep < htol, (eq, ep)
    assert gq < gtol and gp < gtol, (gq, gp)
    c.close()


@pytest.mark.parametrize("form", ["f32", "bf16"])
def test_mixed_converges_to_f64_optimum(oa, O, form):
    print("mixed convergence", form, flush=True)
    I, J, K = 1200, 6000, 4
    res = {}
    for mode in ("f64", form):
        c, d = _ctx(oa, O, I, J, K, seed=77)
        c.set_precision(mode)
        oa.init_em(d, None, 5, d_cu=c, log=None)
        hist = oa.admixture_qn(d, None, 200, 1e-10, d_cu=c, log=None)
        res[mode] = (hist[-1], d.q.copy(), d.p.copy(), len(hist) - 1)
        c.close()
    llf, ll = res["f64"][0], res[form][0]
    dq = np.abs(res[form][1] - res["f64"][1]).max()
    dp = np.abs(res[form][2] - res["f64"][2]).max()
    print(f"{form}: dLL {abs(ll - llf) / abs(llf):.2e}, max|dQ| {dq:.2e}, max|dP| {dp:.2e}, "
          f"iterations {res[form][3]} vs {res['f64'][3]}", flush=True)
    assert abs(ll - llf) <= 1e-6 * abs(llf)


@pytest.mark.parametrize("qsum", [1.0, 0.97])
@pytest.mark.parametrize("I,J,K", [(1, 1, 1), (37, 45, 3), (1500, 4100, 8), (700, 333, 12), (300, 260, 16)])
def test_matrix_core_loglikelihood(oa, O, I, J, K, qsum):
    """BF16 form: the device state's LL with s and 1 - s from the matrix cores
    (three-term BF16 splits) against the FP64 oracle.  Each term's s and t
    carry a relative error of a few 1e-7 (FP32-level), so the bar is 1e-6
    relative to the LL; over many genotypes the errors average out (a
    6-million-genotype cohort lands near 1e-10).  Columns of q summing to
    0.97 (as after EM steps with missing entries) keep t = 1 - s.  The
    tile-loop entry point stays on the FP64 loop (1e-10)."""
    print("mfma ll", I, J, K, qsum, flush=True)
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, alpha=0.3, miss=0.03, seed=I + K, p_lo=0.01, p_hi=0.99, beta=True)
    rng = np.random.default_rng(K)
    q = rng.dirichlet(0.5 * np.ones(K), size=I).T * qsum
    p = rng.uniform(1e-5, 1 - 1e-5, size=(K, J))
    p[:, : J // 4] = np.clip(1 - 10 ** rng.uniform(-5, -2, size=(K, J // 4)), 1e-5, 1 - 1e-5)  # s close to 1
    q, p = np.asfortranarray(q), np.asfortranarray(p)
    c = oa.HIPAdmixData(I, J, K, 3)
    c.upload_bed(bed.reshape((pitch, J), order="F"))
    c.set_precision("bf16")
    c.set_x(0, np.asfortranarray(np.hstack([q, p])))
    got = c.loglikelihood(0)
    ref = O.loglikelihood(bed, pitch, q, p)
    rel = abs(got - ref) / abs(ref)
    print("relative error", rel, flush=True)
    assert rel <= 1e-6, (got, ref, rel)
    loop = c.loglikelihood_loop(q, p, 0, I, 0, J)
    assert abs(loop - ref) <= 1e-10 * abs(ref)
    c.set_precision("f64")
    assert abs(c.loglikelihood(0) - ref) <= 1e-10 * abs(ref)
    c.close()


@pytest.mark.parametrize("qsum", [1.0, 0.97])
@pytest.mark.parametrize("I,J,K", [(1, 1, 1), (37, 45, 3), (1500, 4100, 8), (700, 333, 12), (300, 260, 16),
                                   (130, 700, 2)])
def test_matrix_core_em_step(oa, O, I, J, K, qsum):
    """BF16 form: one FRAPPE EM step (em_q!, em_p!) with s, t and the sums over
    SNPs / individuals on the matrix cores against the FP64 oracle on the same
    state.  The weights are FP32 and every product pairs two-term BF16 splits
    (about 2^-17 relative per term), so the bar is 2e-5 relative to each
    column's maximum for q_next and 2e-5 absolute for p_next (the ratio
    p_tmp / (p_tmp + p_next) lies in [0, 1]).  Columns of q summing to 0.97 (as
    after EM steps with missing entries) exercise the t = 1 - s correction."""
    print("mfma em", I, J, K, qsum, flush=True)
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, alpha=0.3, miss=0.03, seed=I + K, p_lo=0.01, p_hi=0.99, beta=True)
    rng = np.random.default_rng(K + 5)
    q = rng.dirichlet(0.5 * np.ones(K), size=I).T * qsum
    p = rng.uniform(1e-3, 1 - 1e-3, size=(K, J))
    p[:, : J // 4] = np.clip(1 - 10 ** rng.uniform(-3, -1, size=(K, J // 4)), 1e-3, 1 - 1e-3)  # s close to 1
    x = np.asfortranarray(np.hstack([q, p]))
    c = oa.HIPAdmixData(I, J, K, 3)
    c.upload_bed(bed.reshape((pitch, J), order="F"))
    c.set_precision("bf16")
    c.set_x(0, x)
    c.em_q()
    c.em_p()
    got = c.get_x(1)
    st = O.OracleAdmix(bed, pitch, I, J, K, 3, x)
    st.em_q()
    st.em_p()
    ref = st.x_next
    eq = (np.abs(got[:, :I] - ref[:, :I]).max(0) / np.abs(ref[:, :I]).max(0)).max()
    ep = np.abs(got[:, I:] - ref[:, I:]).max()
    print(f"q_next {eq:.2e} relative, p_next {ep:.2e} absolute", flush=True)
    assert eq <= 2e-5 and ep <= 2e-5, (eq, ep)
    # the FP64 form on the same context still matches the oracle to 1e-10
    c.set_precision("f64")
    c.em_q()
    c.em_p()
    f64 = c.get_x(1)
    assert np.abs(f64[:, :I] - ref[:, :I]).max() <= 1e-10 * np.abs(ref[:, :I]).max()
    c.close()


def test_matrix_core_em_priming_lls(oa, O):
    """init_em! in the BF16 form (matrix-core EM sums and log-likelihood) follows
    the FP64 oracle's per-step log-likelihoods to 1e-6 relative."""
    print("mfma init_em", flush=True)
    I, J, K = 1200, 3000, 5
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, alpha=0.3, miss=0.02, seed=11)
    rng = np.random.default_rng(3)
    x = rng.random((K, I + J))
    x[:, :I] /= x[:, :I].sum(0)
    x = np.asfortranarray(x)
    c = oa.HIPAdmixData(I, J, K, 3)
    c.upload_bed(bed.reshape((pitch, J), order="F"))
    c.set_precision("bf16")
    c.set_x(0, x)
    got = c.init_em(5)
    st = O.OracleAdmix(bed, pitch, I, J, K, 3, x)
    ref = st.init_em(5)
    rel = np.abs(got - ref) / np.abs(ref)
    print("per-step LL relative error", rel, flush=True)
    assert rel.max() <= 1e-6
    c.close()
