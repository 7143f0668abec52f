"""Mixed-precision forms of the device-resident update passes (FP64 context,
FP32 or BF16-MFMA passes, FP64 QP / log-likelihood / convergence test).

Per-output error bars against the FP64 oracle on the same state, written
per test:
  Hessian (XtX)   FP32: 1e-5, BF16: 5e-3   relative to the column maximum
  gradient (Xtz)  FP32: 1e-5, BF16: 1e-5   relative to the column maximum
  converged LL    |LL_form - LL_oracle| / |LL_oracle| < 1e-6 (the benchmark bar is 1e-4)
  converged Q, P  FP32: 2e-6, BF16: 5e-5 max-abs (the plain FP64 bar is 1e-6)
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


@pytest.mark.parametrize("K", [2, 3, 8, 12])
@pytest.mark.parametrize("form", ["f32", "bf16"])
def test_mixed_hessians_match_oracle(oa, O, K, form):
    """XtX / Xtz of update_q! and update_p! in each form against the FP64
    oracle's update_q! / update_p! (src/algorithms_inner.jl:167-310) on the same
    state; the plain FP64 passes of the same context meet 1e-9."""
    print("mixed hessians", K, form, flush=True)
    I, J = 1500, 4100
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, alpha=0.3, miss=0.02, seed=K, p_lo=0.01, p_hi=0.99, beta=True)
    c = oa.HIPAdmixData(I, J, K, 3)
    c.upload_bed(bed.reshape((pitch, J), order="F"))
    d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(K))
    oa.init_em(d, None, 3, d_cu=c, log=None)
    ref = O.OracleAdmix(bed, pitch, I, J, K, 3, d.x)
    ref.update_q(False)
    ref.x_next[:, :I] = d.x[:, :I]  # update_p! reads q_next: the same q for every form
    ref.update_p(False)
    out = {}
    for mode in ("f64", form):
        c.set_precision(mode)
        c.update_q(False)
        XtXq, Xtzq = c.hessians(0)
        c.set_x(1, d.x)
        c.update_p(False)
        XtXp, Xtzp = c.hessians(1)
        out[mode] = (XtXq, Xtzq, XtXp, Xtzp)
    htol, gtol = TOL[form]
    for mode, (ht, gt) in (("f64", (1e-9, 1e-9)), (form, (htol, gtol))):
        m = out[mode]
        eq, gq = _colrel(m[0], ref.XtX_q), _colrel(m[1], ref.Xtz_q)
        ep, gp = _colrel(m[2], ref.XtX_p), _colrel(m[3], ref.Xtz_p)
        print(f"K={K} {mode} vs oracle: XtX_q {eq:.2e} Xtz_q {gq:.2e} XtX_p {ep:.2e} Xtz_p {gp:.2e}", flush=True)
        assert eq < ht and ep < ht, (mode, eq, ep)
        assert gq < gt and gp < gt, (mode, gq, gp)
    c.close()


@pytest.mark.parametrize("I,J,K", [(1, 1, 1), (37, 45, 3), (129, 70, 5), (1500, 4100, 8), (700, 333, 12),
                                   (300, 260, 16), (1100, 130, 7)])
def test_matrix_core_update_passes(oa, O, I, J, K):
    """BF16 form's update_q! and update_p! with s, the Hessian and the gradient
    on the matrix cores (ADMX_TUNE_QSIDE_MFMA / _PSIDE_MFMA = 1) and with the
    Hessian only (0), against the FP64 oracle's update_q! / update_p! on the
    same state: Hessian 5e-3, gradient 1e-5 relative to the column maximum.
    Ragged I and J leave partial 128-individual (128-SNP) workgroups and
    partial 32-SNP (32-individual) blocks; a quarter of the SNPs have p close
    to 1 (t = 1 - s down to 1e-5), and q's columns sum to 0.98."""
    from openadmixture_jl_amd import _native as N

    print("matrix-core update passes", I, J, K, flush=True)
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, alpha=0.3, miss=0.03, seed=I + J + K, p_lo=0.01, p_hi=0.99,
                                     beta=True)
    rng = np.random.default_rng(K + 1)
    q = rng.dirichlet(0.5 * np.ones(K), size=I).T * 0.98
    p = rng.uniform(1e-5, 1 - 1e-5, size=(K, J))
    p[:, : J // 4] = np.clip(1 - 10 ** rng.uniform(-5, -2, size=(K, J // 4)), 1e-5, 1 - 1e-5)
    x = np.asfortranarray(np.hstack([q, p]))
    ref = O.OracleAdmix(bed, pitch, I, J, K, 3, x)
    ref.update_q(False)
    ref.x_next[:, :I] = q  # update_p! reads q_next
    ref.update_p(False)
    c = oa.HIPAdmixData(I, J, K, 3)
    c.upload_bed(bed.reshape((pitch, J), order="F"))
    c.set_precision("bf16")
    htol, gtol = TOL["bf16"]
    for mc in (1, 0):
        c.set_tuning(N.ADMX_TUNE_QSIDE_MFMA, mc)
        c.set_tuning(N.ADMX_TUNE_PSIDE_MFMA, mc)
        c.set_x(0, x)
        c.update_q(False)
        XtXq, Xtzq = c.hessians(0)
        c.set_x(1, x)
        c.update_p(False)
        XtXp, Xtzp = c.hessians(1)
        eq, gq = _colrel(XtXq, ref.XtX_q), _colrel(Xtzq, ref.Xtz_q)
        ep, gp = _colrel(XtXp, ref.XtX_p), _colrel(Xtzp, ref.Xtz_p)
        print(f"  {'matrix cores' if mc else 'vector ALU gradient'}: XtX_q {eq:.2e} Xtz_q {gq:.2e} "
              f"XtX_p {ep:.2e} Xtz_p {gp:.2e}", flush=True)
        for a in (XtXq, Xtzq, XtXp, Xtzp):
            assert np.isfinite(a).all()
        assert eq < htol and gq < gtol and ep < htol and gp < gtol, (mc, eq, gq, ep, gp)
    c.close()


CONVERGED_QP_TOL = {"f32": 2e-6, "bf16": 5e-5}


@pytest.mark.parametrize("form", ["f32", "bf16"])
def test_mixed_converges_to_oracle_optimum(oa, O, form):
    """Each form from the same random start and EM(5) priming, to rtol 1e-10,
    against the FP64 oracle run to the same rtol: final LL within 1e-6
    relative, Q and P within the form's bar (FP32 2e-6, BF16 5e-5 max-abs)."""
    print("mixed convergence", form, flush=True)
    I, J, K = 1000, 4000, 4
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, alpha=0.3, miss=0.02, seed=77, p_lo=0.01, p_hi=0.99, beta=True)
    d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(77))
    ref = O.OracleAdmix(bed, pitch, I, J, K, 3, d.x)
    ref.init_em(5)
    rits = 0
    for it in range(1, 301):
        rits = it
        if ref.qn_iteration(it)["reldiff"] < 1e-10:
            break
    c = oa.HIPAdmixData(I, J, K, 3)
    c.upload_bed(bed.reshape((pitch, J), order="F"))
    c.set_precision(form)
    oa.init_em(d, None, 5, d_cu=c, log=None)
    hist = oa.admixture_qn(d, None, 300, 1e-10, d_cu=c, log=None)
    c.close()
    ll, llr = hist[-1], ref.ll_new
    dq = np.abs(d.q - ref.x[:, :I]).max()
    dp = np.abs(d.p - ref.x[:, I:]).max()
    print(f"{form} vs oracle: dLL {abs(ll - llr) / abs(llr):.2e}, max|dQ| {dq:.2e}, max|dP| {dp:.2e}, "
          f"iterations {len(hist) - 1} vs {rits}", flush=True)
    assert abs(ll - llr) <= 1e-6 * abs(llr)
    assert dq <= CONVERGED_QP_TOL[form] and dp <= CONVERGED_QP_TOL[form], (dq, dp)


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
    st = O.OracleAdmix(bed, pitch, I, J, K, 3, x)
    st.em_q()
    st.em_p()
    ref = st.x_next
    from openadmixture_jl_amd import _native as N

    # em_q on the block-image kernel (default) and on k_emq_mfma (QSIDE_MFMA = 0)
    for mc in (-1, 0):
        c.set_tuning(N.ADMX_TUNE_QSIDE_MFMA, mc)
        c.set_x(0, x)
        c.em_q()
        c.em_p()
        got = c.get_x(1)
        eq = (np.abs(got[:, :I] - ref[:, :I]).max(0) / np.abs(ref[:, :I]).max(0)).max()
        ep = np.abs(got[:, I:] - ref[:, I:]).max()
        print(f"em_q kernel {'block images' if mc else 'k_emq_mfma'}: q_next {eq:.2e} relative, "
              f"p_next {ep:.2e} absolute", flush=True)
        assert eq <= 2e-5 and ep <= 2e-5, (mc, eq, ep)
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


def _ll_class(v):
    return "nan" if np.isnan(v) else ("-inf" if v == -np.inf else "finite")


def test_bf16_loglikelihood_degenerate_states(oa, O):
    """BF16 form: states where some s = q.p is exactly 0 or 1 (or p is NaN)
    are flagged by the state check and summed term by term in FP64, so the
    matrix-core log-likelihood gives the reference's NaN / -Inf
    (src/loops.jl:517-539) where the oracle does; a safe state stays on the
    matrix cores (1e-6), a flagged finite one matches to 1e-10."""
    print("bf16 degenerate LL", flush=True)
    I, J, K = 300, 64, 3
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.02, seed=61)
    B = bed.reshape((pitch, J), order="F").copy()
    B[: I // 4, 20] = 0xAA  # SNP 20: every individual heterozygous (code 10), none missing
    B[: I // 4, 21] = 0x00  # SNP 21: monomorphic, every genotype 0
    bed2 = np.asfortranarray(B).ravel(order="F")
    c = oa.HIPAdmixData(I, J, K, 3)
    c.upload_bed(bed2.reshape((pitch, J), order="F"))
    c.set_precision("bf16")
    rng = np.random.default_rng(7)
    q = rng.random((K, I)) + 0.05
    q = np.asfortranarray(q / q.sum(0))
    p = np.asfortranarray(rng.uniform(0.02, 0.98, size=(K, J)))
    cases = {"plain": (p, 1e-6)}
    for name, (col, val) in {"zero_p_mixed_snp": (5, 0.0), "zero_p_het_snp": (20, 0.0),
                             "zero_p_mono_snp": (21, 0.0), "one_p": (7, 1.0),
                             "nan_p": (9, np.nan), "tiny_p": (11, 1e-13)}.items():
        pp = p.copy()
        pp[:, col] = val
        cases[name] = (np.asfortranarray(pp), 1e-10)
    seen = set()
    for name, (pp, tol) in cases.items():
        ref = O.loglikelihood(bed2, pitch, q, pp)
        c.set_x(0, np.asfortranarray(np.hstack([q, pp])))
        dev = c.loglikelihood(0)
        print(name, ref, dev, flush=True)
        assert _ll_class(dev) == _ll_class(ref), (name, dev, ref)
        if np.isfinite(ref):
            assert abs(dev - ref) <= tol * abs(ref), (name, dev, ref)
        seen.add(_ll_class(ref))
    assert seen == {"nan", "-inf", "finite"}
    c.close()


def test_bf16_em_on_a_monomorphic_snp_matches_reference(oa, O):
    """BF16 form, FRAPPE EM from a random start on a cohort with a SNP whose
    genotypes are all 0: after the first step p of that SNP is exactly 0; the
    state check sends the following EM steps and log-likelihoods to the FP64
    passes, so the priming log-likelihoods are NaN where the oracle's are."""
    print("bf16 EM monomorphic", flush=True)
    I, J, K = 200, 300, 3
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.0, seed=62)
    B = bed.reshape((pitch, J), order="F").copy()
    B[: I // 4, 40] = 0x00
    bed2 = np.asfortranarray(B).ravel(order="F")
    d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(5))
    c = oa.HIPAdmixData(I, J, K, 3)
    c.upload_bed(bed2.reshape((pitch, J), order="F"))
    c.set_precision("bf16")
    ref = O.OracleAdmix(bed2, pitch, I, J, K, 3, d.x)
    lls = oa.init_em(d, None, 3, d_cu=c, log=None)
    rlls = ref.init_em(3)
    print(lls, rlls, flush=True)
    assert [_ll_class(v) for v in lls] == [_ll_class(v) for v in rlls]
    assert _ll_class(rlls[0]) == "nan"
    assert np.array_equal(np.isnan(d.x), np.isnan(ref.x))
    c.close()
