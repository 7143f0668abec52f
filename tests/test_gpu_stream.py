"""J-chunk streaming (BED in pinned host memory, streamed through HBM per
pass): every loop and the device-resident algorithm agree with the resident
context and with the oracle, for chunk sizes that do and do not divide J."""
import numpy as np
import pytest

from conftest import make_cohort, random_state

pytestmark = pytest.mark.gpu


def _colrel(got, ref):
    got = np.asarray(got).reshape(-1, np.asarray(got).shape[-1])
    ref = np.asarray(ref).reshape(-1, np.asarray(ref).shape[-1])
    return (np.abs(got - ref).max(0) / (np.abs(ref).max(0) + 1e-300)).max()


@pytest.mark.parametrize("S", [16, 100, 1000, 5000])
def test_streamed_loops_match_resident_and_oracle(oa, O, S):
    print("stream loops", S, flush=True)
    I, J, K = 500, 2100, 4
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.03, seed=51)
    res = oa.HIPAdmixData(I, J, K, 3)
    st = oa.HIPAdmixData(I, J, K, 3, stream_snps=S)
    for c in (res, st):
        c.upload_bed(bed.reshape((pitch, J), order="F"))
    assert np.array_equal(st.download_bed().ravel(order="F"), bed)
    q, p = random_state(I, J, K, 3)
    for (i0, i1, j0, j1) in [(0, I, 0, J), (5, I - 3, 17, J - 9)]:
        ll_s = st.loglikelihood_loop(q, p, i0, i1, j0, j1)
        ll_r = res.loglikelihood_loop(q, p, i0, i1, j0, j1)
        ll_o = O.loglikelihood(bed, pitch, q, p, (i0, i1), (j0, j1))
        assert abs(ll_s - ll_r) <= 1e-12 * abs(ll_r) and abs(ll_s - ll_o) <= 1e-10 * abs(ll_o)
        out = {}
        for name, c in (("s", st), ("r", res)):
            XQ, ZQ = np.zeros((K, K, I), order="F"), np.zeros((K, I), order="F")
            XP, ZP = np.zeros((K, K, J), order="F"), np.zeros((K, J), order="F")
            c.update_q_loop(XQ, ZQ, q, p, i0, i1, j0, j1)
            c.update_p_loop(XP, ZP, q, p, i0, i1, j0, j1)
            pn, pt = np.zeros((K, J), order="F"), np.zeros((K, J), order="F")
            c.em_p_loop(pn, pt, q, p, i0, i1, j0, j1)
            qn = np.zeros((K, I), order="F")
            c.em_q_loop(qn, q, p, i0, i1, j0, j1)
            out[name] = (XQ, ZQ, XP, ZP, pn, pt, qn)
        for a, b in zip(out["s"], out["r"]):
            assert _colrel(a, b) < 1e-12
        # the P side is chunk-local: bit-identical
        assert np.array_equal(out["s"][2], out["r"][2]) and np.array_equal(out["s"][3], out["r"][3])
    with pytest.raises(oa.AdmixHIPError):
        st.decode()
    res.close()
    st.close()


def test_streamed_algorithm_matches_oracle(oa, O):
    print("stream algorithm", flush=True)
    I, J, K = 600, 3000, 3
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.02, seed=52)
    d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(9))
    c = oa.HIPAdmixData(I, J, K, 3, stream_snps=700)
    c.upload_bed(bed.reshape((pitch, J), order="F"))
    ref = O.OracleAdmix(bed, pitch, I, J, K, 3, d.x, d.U, d.V)
    np.testing.assert_allclose(oa.init_em(d, None, 3, d_cu=c, log=None), ref.init_em(3), rtol=1e-10, atol=0)
    for it in range(1, 6):
        r = c.qn_iteration(it)
        rr = ref.qn_iteration(it)
        for key in ("basic", "qn", "new"):
            assert abs(r[key] - rr[key]) <= 1e-10 * abs(rr[key]), (it, key)
    c.close()


def test_streamed_generator_matches_resident(oa, O):
    I, J, K = 300, 1000, 3
    rng = np.random.default_rng(2)
    Qt = np.asfortranarray(rng.dirichlet(np.ones(K), size=I).T)
    Pt = np.asfortranarray(rng.uniform(0.05, 0.95, (K, J)))
    a = oa.HIPAdmixData(I, J, K, 3)
    b = oa.HIPAdmixData(I, J, K, 3, stream_snps=333)
    a.generate_bed(Qt, Pt, 5, 0.01)
    b.generate_bed(Qt, Pt, 5, 0.01)
    assert np.array_equal(a.download_bed(), b.download_bed())
    a.close()
    b.close()


def _plink(oa, tmp_path, bed, pitch, I, J):
    """Write the cohort as a PLINK trio; returns (.bed path, raw payload)."""
    data = np.asfortranarray(bed.reshape((pitch, J), order="F")[: (I + 3) // 4])
    return oa.write_plink(str(tmp_path / "cohort"), data, I, J), data


# (chunk SNPs, SNPs kept in HBM): none, part, all (R = J), more than J (clamped),
# as many as fit (-1: all of this small cohort)
@pytest.mark.parametrize("S,R", [(16, 0), (700, 0), (700, 1000), (333, 2100), (1000, 5000), (100, 37), (100, -1)])
def test_file_streamed_loops_match_resident(oa, O, tmp_path, S, R):
    """A context streaming straight from the .bed file (pread into pinned
    staging, re-pitched in HBM), with SNPs [0, R) resident: every loop matches
    the resident context (P side bit-identical) and the oracle."""
    print("file stream loops", S, R, flush=True)
    I, J, K = 501, 2100, 4
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.03, seed=53)
    path, data = _plink(oa, tmp_path, bed, pitch, I, J)
    res = oa.HIPAdmixData(I, J, K, 3)
    res.upload_bed(bed.reshape((pitch, J), order="F"))
    st = oa.HIPAdmixData(I, J, K, 3, stream_snps=S, bed_file=path, resident_snps=R)
    assert np.array_equal(st.download_bed(), res.download_bed())
    assert np.array_equal(st.download_bed(), data)
    q, p = random_state(I, J, K, 4)
    for (i0, i1, j0, j1) in [(0, I, 0, J), (5, I - 3, 17, J - 9), (0, I, 990, 1010)]:
        ll_s = st.loglikelihood_loop(q, p, i0, i1, j0, j1)
        ll_r = res.loglikelihood_loop(q, p, i0, i1, j0, j1)
        ll_o = O.loglikelihood(bed, pitch, q, p, (i0, i1), (j0, j1))
        assert abs(ll_s - ll_r) <= 1e-12 * abs(ll_r) and abs(ll_s - ll_o) <= 1e-10 * abs(ll_o)
        out = {}
        for name, c in (("s", st), ("r", res)):
            XQ, ZQ = np.zeros((K, K, I), order="F"), np.zeros((K, I), order="F")
            XP, ZP = np.zeros((K, K, J), order="F"), np.zeros((K, J), order="F")
            c.update_q_loop(XQ, ZQ, q, p, i0, i1, j0, j1)
            c.update_p_loop(XP, ZP, q, p, i0, i1, j0, j1)
            qn = np.zeros((K, I), order="F")
            c.em_q_loop(qn, q, p, i0, i1, j0, j1)
            out[name] = (XQ, ZQ, XP, ZP, qn)
        for a, b in zip(out["s"], out["r"]):
            assert _colrel(a, b) < 1e-12
        assert np.array_equal(out["s"][2], out["r"][2]) and np.array_equal(out["s"][3], out["r"][3])
    with pytest.raises(oa.AdmixHIPError, match="file-backed"):
        st.upload_bed(data)
    res.close()
    st.close()


@pytest.mark.parametrize("source,R", [("pinned", 1234), ("file", 0), ("file", 1234)])
def test_streamed_sources_algorithm_matches_oracle(oa, O, tmp_path, source, R):
    """init_em + admixture_qn! iterations on a pinned or file-backed stream
    with part of the SNPs resident in HBM, against the oracle."""
    print("stream algorithm", source, R, flush=True)
    I, J, K = 600, 3000, 3
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.02, seed=52)
    d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(9))
    if source == "file":
        path, _ = _plink(oa, tmp_path, bed, pitch, I, J)
        c = oa.HIPAdmixData(I, J, K, 3, stream_snps=700, bed_file=path, resident_snps=R)
    else:
        c = oa.HIPAdmixData(I, J, K, 3, stream_snps=700)
        c.set_resident_snps(R)  # before the upload: refreshed by it
        c.upload_bed(bed.reshape((pitch, J), order="F"))
    ref = O.OracleAdmix(bed, pitch, I, J, K, 3, d.x, d.U, d.V)
    np.testing.assert_allclose(oa.init_em(d, None, 3, d_cu=c, log=None), ref.init_em(3), rtol=1e-10, atol=0)
    for it in range(1, 6):
        r = c.qn_iteration(it)
        rr = ref.qn_iteration(it)
        for key in ("basic", "qn", "new"):
            assert abs(r[key] - rr[key]) <= 1e-10 * abs(rr[key]), (it, key)
    c.close()


def test_resident_snps_follow_the_genotypes(oa, O):
    """Pinned stream: the HBM-resident SNPs are refreshed by a new upload and
    by the generator, and the count can be changed or cleared at any time."""
    print("resident refresh", flush=True)
    I, J, K = 300, 1000, 3
    q, p = random_state(I, J, K, 5)
    a = oa.HIPAdmixData(I, J, K, 3)
    b = oa.HIPAdmixData(I, J, K, 3, stream_snps=128, resident_snps=500)
    for seed in (54, 55):
        _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.02, seed=seed)
        for c in (a, b):
            c.upload_bed(bed.reshape((pitch, J), order="F"))
        ref = O.loglikelihood(bed, pitch, q, p)
        for R in (500, 999, 0, 1000):
            b.set_resident_snps(R)
            assert abs(b.loglikelihood_loop(q, p, 0, I, 0, J) - ref) <= 1e-10 * abs(ref)
    Qt = np.asfortranarray(np.random.default_rng(3).dirichlet(np.ones(K), size=I).T)
    Pt = np.asfortranarray(np.random.default_rng(4).uniform(0.05, 0.95, (K, J)))
    for c in (a, b):
        c.generate_bed(Qt, Pt, 7, 0.01)
    assert abs(b.loglikelihood_loop(q, p, 0, I, 0, J) - a.loglikelihood_loop(q, p, 0, I, 0, J)) <= 1e-12 * abs(
        a.loglikelihood_loop(q, p, 0, I, 0, J))
    with pytest.raises(oa.AdmixHIPError):
        a.set_resident_snps(10)  # resident context
    with pytest.raises(oa.AdmixHIPError):
        b.set_resident_snps(-2)
    a.close()
    b.close()


def test_run_admixture_streams_from_the_file(oa, O, tmp_path):
    """run_admixture(...; stream_snps, stream_from_file, resident_snps) fits
    from the .bed file without a host copy and reaches the resident fit's LL."""
    print("run_admixture file stream", flush=True)
    I, J, K = 400, 1500, 3
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.01, seed=56)
    path, _ = _plink(oa, tmp_path, bed, pitch, I, J)
    fits = {}
    for name, kw in (("resident", {}), ("file", dict(stream_snps=256, stream_from_file=True, resident_snps=700))):
        d, _, _ = oa.run_admixture(path, K, rng=np.random.default_rng(4), admix_n_iter=40, log=None, **kw)
        fits[name] = d.ll_new
        if name == "file":
            assert d.hip.bed_file == path and d.hip.stream_snps == 256
        d.hip.close()
    assert abs(fits["file"] - fits["resident"]) <= 1e-8 * abs(fits["resident"])
