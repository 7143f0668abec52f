"""GPU parity: every entry point of libadmixhip, called through the C ABI,
against the CPU oracle on the same seeded inputs.

Tolerances (written per test): decode / counts / generator bit-exact;
log-likelihood 1e-10 relative in Float64 and 1e-5 in Float32; gradients and
Hessians 1e-9 (F64) / 2e-4 (F32) relative to the column's largest entry;
QP steps and projected iterates 1e-9 max-abs in Float64.
"""
import numpy as np
import pytest

from conftest import make_cohort, random_state

pytestmark = pytest.mark.gpu

F64, F32 = np.float64, np.float32


def note(*a):
    print(*a, flush=True)


def ctx_for(oa, bed, pitch, I, J, K, Q=3, T=F64):
    c = oa.HIPAdmixData(I, J, K, Q, T)
    c.upload_bed(np.asarray(bed).reshape((pitch, J), order="F"))
    return c


# ------------------------------------------------------------------ genotypes
@pytest.mark.parametrize("I,J", [(1, 1), (13, 7), (1000, 300), (4099, 257)])
def test_decode_counts_bit_exact(oa, O, I, J):
    note("decode/counts", I, J)
    Qt, Pt, bed, pitch = make_cohort(O, I, J, 3, miss=0.05, seed=I + J)
    c = ctx_for(oa, bed, pitch, I, J, 3)
    g = c.decode()
    assert np.array_equal(g, O.decode(bed, pitch, I, J))
    snm, salt, inm = c.counts()
    rs, ra, ri = O.counts(bed, pitch, I, J)
    assert np.array_equal(snm, rs) and np.array_equal(salt, ra) and np.array_equal(inm, ri)
    # round trip of the payload through HBM (pitch-padded on device)
    assert np.array_equal(c.download_bed().reshape(-1, order="F"), bed)
    c.close()


def test_decode_known_bytes(oa):
    # 0xE4 = 11 10 01 00 (low bits first): individuals 0..3 -> 0, missing, 1, 2
    I, J = 4, 2
    bed = np.array([[0xE4, 0x1B]], dtype=np.uint8).reshape((1, 2), order="F")
    c = oa.HIPAdmixData(I, J, 2)
    c.upload_bed(bed)
    g = c.decode()
    assert g[:, 0].tolist() == [0, -1, 1, 2]
    assert g[:, 1].tolist() == [2, 1, -1, 0]
    snm, salt, inm = c.counts()
    assert snm.tolist() == [3, 3] and salt.tolist() == [3, 3] and inm.tolist() == [2, 1, 1, 2]
    c.close()


@pytest.mark.parametrize("I,J,K", [(999, 500, 3), (64, 64, 8)])
def test_generator_matches_oracle(oa, O, I, J, K):
    note("generator", I, J, K)
    rng = np.random.default_rng(4)
    Qt = np.asfortranarray(rng.dirichlet(0.5 * np.ones(K), size=I).T)
    Pt = np.asfortranarray(rng.uniform(0.05, 0.95, size=(K, J)))
    c = oa.HIPAdmixData(I, J, K)
    c.generate_bed(Qt, Pt, 77, 0.02)
    ref = O.gen_bed(Qt, Pt, 77, 0.02)
    assert np.array_equal(c.download_bed().reshape(-1, order="F"), ref)
    c.close()


# ------------------------------------------------------------------ tile loops
CASES = [(700, 3000, 4, 0), (333, 1025, 3, 1), (257, 515, 8, 2), (100, 200, 1, 3), (130, 300, 12, 4), (40, 90, 16, 5)]


@pytest.mark.parametrize("T", [F64, F32])
@pytest.mark.parametrize("I,J,K,seed", CASES)
def test_loglikelihood_loop(oa, O, I, J, K, seed, T):
    note("ll loop", I, J, K, T)
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, seed=seed)
    q, p = random_state(I, J, K, seed, T)
    c = ctx_for(oa, bed, pitch, I, J, K, T=T)
    rtol = 1e-10 if T == F64 else 1e-5
    for (i0, i1, j0, j1) in [(0, I, 0, J), (3, I - 5, 7, J - 2), (I // 2, I // 2 + 1, 0, J)]:
        got = c.loglikelihood_loop(q, p, i0, i1, j0, j1)
        ref = O.loglikelihood(bed, pitch, q, p, (i0, i1), (j0, j1))
        assert abs(got - ref) <= rtol * abs(ref), (i0, i1, j0, j1, got, ref)
    assert c.loglikelihood_loop(q, p, 5, 5, 0, J) == 0.0
    c.close()


def _colmax_close(got, ref, rtol):
    got, ref = np.asarray(got, np.float64), np.asarray(ref, np.float64)
    scale = np.abs(ref).reshape(-1, ref.shape[-1]).max(axis=0) + 1e-300
    err = np.abs(got - ref).reshape(-1, ref.shape[-1]).max(axis=0)
    return np.all(err <= rtol * scale), (err / scale).max()


@pytest.mark.parametrize("T", [F64, F32])
@pytest.mark.parametrize("I,J,K,seed", CASES)
def test_update_loops(oa, O, I, J, K, seed, T):
    note("update loops", I, J, K, T)
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, seed=seed)
    q, p = random_state(I, J, K, seed, T)
    c = ctx_for(oa, bed, pitch, I, J, K, T=T)
    rtol = 1e-9 if T == F64 else 2e-4
    rng = np.random.default_rng(seed)
    i0, i1, j0, j1 = 2, I - 1, 5, J - 3
    # Q side: accumulate into non-zero outputs over a sub-range
    XtX = np.asfortranarray(rng.random((K, K, I)).astype(T))
    Xtz = np.asfortranarray(rng.random((K, I)).astype(T))
    gX, gZ = XtX.copy(order="F"), Xtz.copy(order="F")
    c.update_q_loop(gX, gZ, q, p, i0, i1, j0, j1)
    rX, rZ = XtX.copy(order="F"), Xtz.copy(order="F")
    O.update_q_loop(rX, rZ, bed, pitch, q, p, (i0, i1), (j0, j1))
    ok, e = _colmax_close(gX, rX, rtol)
    assert ok, e
    ok, e = _colmax_close(gZ, rZ, rtol)
    assert ok, e
    # P side
    XtX = np.asfortranarray(rng.random((K, K, J)).astype(T))
    Xtz = np.asfortranarray(rng.random((K, J)).astype(T))
    gX, gZ = XtX.copy(order="F"), Xtz.copy(order="F")
    c.update_p_loop(gX, gZ, q, p, i0, i1, j0, j1)
    rX, rZ = XtX.copy(order="F"), Xtz.copy(order="F")
    O.update_p_loop(rX, rZ, bed, pitch, q, p, (i0, i1), (j0, j1))
    ok, e = _colmax_close(gX, rX, rtol)
    assert ok, e
    ok, e = _colmax_close(gZ, rZ, rtol)
    assert ok, e
    c.close()


def test_update_q_gradient_near_the_bounds(oa, O):
    """The FP64 update_q pass forms Xtz = W q - sum b p + sum d from its Hessian
    accumulators (W = sum (a+b) p p').  Where s is within 1e-5 of 1 and a
    genotype is below 2, b = (2-g)/t^2 reaches 1e10 and the terms of size b
    cancel to one of size d (1-p): the difference from the direct two-term sums
    is then bounded by a few unit roundoffs of sum_j b_ij, as for the rank-1
    Hessian.  A quarter of the SNPs sit at the projection's bound 1 - 1e-5 in
    every population (a state the fits never reach; it maximises the
    cancellation)."""
    note("update_q gradient near the bounds")
    I, J, K = 300, 2000, 6
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, seed=21)
    q, p = random_state(I, J, K, seed=21)
    p[:, : J // 4] = 1 - 1e-5
    c = ctx_for(oa, bed, pitch, I, J, K)
    gX = np.zeros((K, K, I), order="F")
    gZ = np.zeros((K, I), order="F")
    c.update_q_loop(gX, gZ, q, p, 0, I, 0, J)
    rX = np.zeros((K, K, I), order="F")
    rZ = np.zeros((K, I), order="F")
    O.update_q_loop(rX, rZ, bed, pitch, q, p, (0, I), (0, J))
    g = O.decode(bed, pitch, I, J).astype(np.float64)
    t = 1.0 - q.T @ p
    b = np.where(g >= 0, 2.0 - g, 0.0) / t**2
    bound = 64 * np.finfo(np.float64).eps * b.sum(1) + 1e-9 * np.abs(rZ).max(0)
    err = np.abs(gZ - rZ).max(0)
    assert np.all(err <= bound), (err / bound).max()
    # the same bound holds the rank-1 Hessian
    errX = np.abs(gX - rX).max((0, 1))
    assert np.all(errX <= bound + 1e-9 * np.abs(rX).max((0, 1))), (errX / bound).max()
    # away from the bound (p in [0.02, 0.98]) the gradient matches the direct sums closely
    j0 = J // 4
    gZ[...] = 0
    rZ[...] = 0
    c.update_q_loop(gX, gZ, q, p, 0, I, j0, J)
    O.update_q_loop(rX, rZ, bed, pitch, q, p, (0, I), (j0, J))
    assert (np.abs(gZ - rZ).max(0) / np.abs(rZ).max(0)).max() < 1e-11
    c.close()


@pytest.mark.parametrize("T", [F64, F32])
@pytest.mark.parametrize("I,J,K,seed", CASES[:3])
def test_em_loops(oa, O, I, J, K, seed, T):
    note("em loops", I, J, K, T)
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, seed=seed)
    q, p = random_state(I, J, K, seed, T)
    c = ctx_for(oa, bed, pitch, I, J, K, T=T)
    rtol = 1e-10 if T == F64 else 1e-4
    i0, i1, j0, j1 = 1, I - 2, 3, J - 1
    gq = np.zeros((K, I), T, order="F")
    rq = np.zeros((K, I), T, order="F")
    c.em_q_loop(gq, q, p, i0, i1, j0, j1)
    O.em_q_loop(rq, bed, pitch, q, p, (i0, i1), (j0, j1))
    ok, e = _colmax_close(gq, rq, rtol)
    assert ok, e
    gpn, gpt = np.zeros((K, J), T, order="F"), np.zeros((K, J), T, order="F")
    rpn, rpt = np.zeros((K, J), T, order="F"), np.zeros((K, J), T, order="F")
    c.em_p_loop(gpn, gpt, q, p, i0, i1, j0, j1)
    O.em_p_loop(rpn, rpt, bed, pitch, q, p, (i0, i1), (j0, j1))
    for a, b in ((gpn, rpn), (gpt, rpt)):
        ok, e = _colmax_close(a, b, rtol)
        assert ok, e
    c.close()


# ------------------------------------------------------------------ QP / projections / QN
@pytest.mark.parametrize("K", [2, 3, 5, 8, 12, 16])
@pytest.mark.parametrize("simplex", [True, False])
def test_qp_batch_matches_oracle(oa, O, K, simplex):
    note("qp batch", K, simplex)
    I, J = 300, 400
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, seed=K)
    q, p = random_state(I, J, K, K)
    # realistic tableaux: the Hessians of an update pass
    if simplex:
        XtX = np.zeros((K, K, I), order="F")
        Xtz = np.zeros((K, I), order="F")
        O.update_q_loop(XtX, Xtz, bed, pitch, q, p)
        x = q
    else:
        XtX = np.zeros((K, K, J), order="F")
        Xtz = np.zeros((K, J), order="F")
        O.update_p_loop(XtX, Xtz, bed, pitch, q, p)
        x = p
    Xtz = np.asfortranarray(-Xtz)
    c = oa.HIPAdmixData(I, J, K)
    delta, status = c.qp_batch(simplex, XtX, Xtz, x)
    n = x.shape[1]
    for col in range(n):
        tab = O.create_tableau(np.asfortranarray(XtX[:, :, col]), Xtz[:, col].copy(), x[:, col].copy(), simplex)
        rd, rs = O.quadratic_program(tab, x[:, col].copy(), np.zeros(K), np.ones(K), 1 if simplex else 0)
        assert status[col] == rs, (col, status[col], rs)
        assert np.abs(delta[:, col] - rd).max() <= 1e-9, (col, delta[:, col], rd)
    c.close()


@pytest.mark.parametrize("K", [1, 2, 3, 7, 16])
def test_projections(oa, O, K):
    note("projections", K)
    rng = np.random.default_rng(K)
    n = 2000
    b = np.asfortranarray(rng.normal(0.3, 0.6, size=(K, n)))
    c = oa.HIPAdmixData(4, 4, K)
    got = c.project_q(b.copy(order="F"))
    ref = O.project_q(b.copy(order="F"))
    assert np.array_equal(got, ref)
    v = rng.normal(0.5, 0.7, size=5000)
    assert np.array_equal(c.project_p(v.copy()), O.project_p(v.copy()))
    c.close()


@pytest.mark.parametrize("it,Q", [(1, 3), (2, 3), (3, 3), (7, 3), (5, 1), (4, 8)])
def test_zal_step(oa, O, it, Q):
    note("zal", it, Q)
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
    c.zal_step(it, x, xn, xn2, gU, gV, gx)
    rU, rV = U0.copy(order="F"), V0.copy(order="F")
    O.update_UV(rU, rV, x, xn, xn2, it, Q)
    rx = np.zeros_like(x, order="F")
    O.update_QN(rx, xn, x, rU, rV, min(it, Q))
    O.project_p(rx[:, I:])
    rxq = np.asfortranarray(rx[:, :I])
    O.project_q(rxq)
    rx[:, :I] = rxq
    assert np.array_equal(gU, rU) and np.array_equal(gV, rV)
    assert np.abs(gx - rx).max() < 1e-9
    c.close()


@pytest.mark.parametrize("K", [5, 9, 10, 12])
def test_update_steps_without_kept_hessians(oa, O, K):
    """With ADMX_TUNE_KEEP_HESSIANS = 0 the split finalize (reduction, then QP)
    keeps its Hessians in the context's own arrays: the steps and QP codes are
    the same bit for bit."""
    note("keep_hess", K)
    I, J = 500, 1500
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, seed=K)
    q, p = random_state(I, J, K, seed=K)
    c = ctx_for(oa, bed, pitch, I, J, K)
    x = np.asfortranarray(np.hstack([q, p]))
    out = []
    for keep in (1, 0):
        c.set_tuning(oa._native.ADMX_TUNE_KEEP_HESSIANS, keep)
        c.set_x(0, x)
        sq = c.update_q(False, status=True)
        sp = c.update_p(False, status=True)
        out.append((c.get_x(1), sq, sp))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])
    c.close()


# ------------------------------------------------------------------ algorithms on device state
def _pair(oa, O, I, J, K, Q=3, T=F64, seed=0, miss=0.01, beta=False, p_lo=0.05, p_hi=0.95):
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, seed=seed, miss=miss, beta=beta, p_lo=p_lo, p_hi=p_hi)
    d = oa.AdmixData(I, J, K, Q, rng=np.random.default_rng(seed + 100), T=T)
    c = ctx_for(oa, bed, pitch, I, J, K, Q, T)
    ref = O.OracleAdmix(bed, pitch, I, J, K, Q, d.x, d.U, d.V, dtype=T)
    return d, c, ref


@pytest.mark.parametrize("K", [4, 8, 10, 12, 13, 16])
def test_update_q_p_device(oa, O, K):
    # K >= 9 (Q side) / 10 (P side) runs the finalize as a reduction kernel
    # followed by the separate QP kernel; K >= 13 splits the Hessian triangle
    # over two pass launches on both sides; K = 16 is the largest K
    note("update_q!/update_p!", K)
    I, J = 600, 2000
    d, c, ref = _pair(oa, O, I, J, K)
    ref.init_em(2)
    d.x[...] = ref.x
    s_q = oa.update_q(d, None, False, d_cu=c)
    r_q = ref.update_q(False)
    assert np.array_equal(s_q, r_q)
    assert np.abs(d.q_next - ref.x_next[:, :I]).max() < 1e-9
    ok, e = _colmax_close(d.XtX_q, ref.XtX_q, 1e-9)
    assert ok, e
    ok, e = _colmax_close(d.Xtz_q, ref.Xtz_q, 1e-9)
    assert ok, e
    d.x_next[:, :I] = ref.x_next[:, :I]
    s_p = oa.update_p(d, None, False, d_cu=c)
    r_p = ref.update_p(False)
    assert np.array_equal(s_p, r_p)
    assert np.abs(d.p_next - ref.x_next[:, I:]).max() < 1e-9
    c.close()


# the last two: a cohort smaller than one wave and one SNP step, and one of
# exactly one wave of individuals and one 16-SNP step (allele frequencies in
# [0.4, 0.6], so that no SNP of these few individuals is monomorphic: see
# test_loglikelihood_degenerate_states for that case)
@pytest.mark.parametrize("I,J,K", [(500, 3000, 3), (1000, 2000, 5), (9, 17, 2), (64, 16, 3)])
def test_em_and_qn_iterations_f64(oa, O, I, J, K):
    note("qn iterations f64", I, J, K)
    tiny = I < 100
    d, c, ref = _pair(oa, O, I, J, K, seed=I, p_lo=0.4 if tiny else 0.05, p_hi=0.6 if tiny else 0.95)
    if tiny:
        g = c.decode()
        assert all(len(set(g[:, j][g[:, j] >= 0].tolist())) > 1 for j in range(J)), "monomorphic SNP"
    lls = oa.init_em(d, None, 5, d_cu=c, log=None)
    rlls = ref.init_em(5)
    np.testing.assert_allclose(lls, rlls, rtol=1e-10, atol=0)
    assert np.abs(d.x - ref.x).max() < 1e-10
    d_ll = []
    for it in range(1, 9):
        r = c.qn_iteration(it)
        rr = ref.qn_iteration(it)
        d_ll.append(r["new"])
        for key in ("basic", "qn", "new"):
            assert abs(r[key] - rr[key]) <= 1e-10 * abs(rr[key]), (it, key, r[key], rr[key])
    x = c.get_x(0)
    assert np.abs(x - ref.x).max() < 1e-6
    assert all(b >= a - 1e-9 * abs(a) for a, b in zip(d_ll, d_ll[1:]))
    c.close()


def test_qn_iterations_f32(oa, O):
    note("qn iterations f32")
    I, J, K = 800, 3000, 3
    d, c, ref = _pair(oa, O, I, J, K, T=F32, seed=9)
    lls = oa.init_em(d, None, 3, d_cu=c, log=None)
    rlls = ref.init_em(3)
    np.testing.assert_allclose(lls, rlls, rtol=1e-5)
    for it in range(1, 4):
        r = c.qn_iteration(it)
        rr = ref.qn_iteration(it)
        assert abs(r["new"] - rr["new"]) <= 1e-5 * abs(rr["new"]), (it, r, rr)
    c.close()


def test_fix_q_fix_p(oa, O):
    note("fix_q/fix_p")
    I, J, K = 300, 900, 3
    d, c, ref = _pair(oa, O, I, J, K, seed=3)
    oa.init_em(d, None, 2, d_cu=c, log=None)
    ref.init_em(2)
    q0 = c.get_x(0)[:, :I].copy()
    r = c.qn_iteration(1, fix_q=True)
    rr = ref.qn_iteration(1, fix_q=True)
    assert abs(r["new"] - rr["new"]) <= 1e-10 * abs(rr["new"])
    if r["qn"] <= r["prev"]:
        assert np.array_equal(c.get_x(0)[:, :I], q0)
    r = c.qn_iteration(2, fix_p=True)
    rr = ref.qn_iteration(2, fix_p=True)
    assert abs(r["new"] - rr["new"]) <= 1e-10 * abs(rr["new"])
    c.close()


def test_reproducible_bitwise(oa, O):
    note("reproducibility")
    I, J, K = 700, 1500, 4
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, seed=21)
    outs = []
    for _ in range(2):
        d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(5))
        c = ctx_for(oa, bed, pitch, I, J, K)
        oa.init_em(d, None, 2, d_cu=c, log=None)
        r = [c.qn_iteration(it)["new"] for it in (1, 2, 3)]
        outs.append((r, c.get_x(0)))
        c.close()
    assert outs[0][0] == outs[1][0]
    assert np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("prec", ["f64", "bf16"])
def test_init_em_ll_new_is_last_step(oa, O, prec):
    """init_em! (src/algorithms_outer.jl:12-45) sums the log-likelihood after
    every EM step and then d.ll_new of the same state; the device keeps the last
    step's sum for ll_new, which must equal a fresh pass over that state bit
    for bit (FP64 passes and the BF16 form's matrix-core pass alike)."""
    note("init_em ll_new", prec)
    I, J, K = 900, 1300, 4
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, seed=23)
    d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(6))
    c = ctx_for(oa, bed, pitch, I, J, K)
    c.set_precision(prec)
    c.set_x(0, d.x)
    lls = c.init_em(3)
    ll_new = c.get_ll()[1]
    assert ll_new == lls[-1]
    assert c.loglikelihood(0) == ll_new
    c.close()



def _ll_class(v):
    return "nan" if np.isnan(v) else ("-inf" if v == -np.inf else "finite")


def test_loglikelihood_degenerate_states(oa, O):
    """States with s = q.p of exactly 0 or 1, as FRAPPE EM produces on a SNP
    whose observed genotypes are all 0 (p -> 0): the reference's terms
    m [g log s + (2-g) log(1-s)] (src/loops.jl:517-539) give NaN for 0 log 0
    (g = 0 or a missing entry at s = 0) and -Inf for g log 0.  The tile loop
    and the device-resident log-likelihood reproduce NaN / -Inf / finite as the
    oracle does, finite values to 1e-10."""
    note("degenerate LL")
    I, J, K = 300, 64, 3
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.02, seed=61)
    B = bed.reshape((pitch, J), order="F").copy()
    B[: I // 4, 20] = 0xAA  # SNP 20: every individual heterozygous (code 10), none missing
    B[: I // 4, 21] = 0x00  # SNP 21: monomorphic, every genotype 0
    bed2 = np.asfortranarray(B).ravel(order="F")
    c = ctx_for(oa, bed2, pitch, I, J, K)
    q, p = random_state(I, J, K, 7)
    cases = {"plain": p}
    for name, (cols, val) in {"zero_p_mixed_snp": ([5], 0.0), "zero_p_het_snp": ([20], 0.0),
                              "zero_p_mono_snp": ([21], 0.0), "one_p": ([7], 1.0), "nan_p": ([9], np.nan),
                              "tiny_p": ([11], 1e-13)}.items():
        pp = p.copy()
        pp[:, cols] = val
        cases[name] = np.asfortranarray(pp)
    seen = set()
    for name, pp in cases.items():
        ref = O.loglikelihood(bed2, pitch, q, pp)
        loop = c.loglikelihood_loop(q, pp, 0, I, 0, J)
        c.set_x(0, np.asfortranarray(np.hstack([q, pp])))
        dev = c.loglikelihood(0)
        print(name, ref, loop, dev, flush=True)
        for got in (loop, dev):
            assert _ll_class(got) == _ll_class(ref), (name, got, ref)
            if np.isfinite(ref):
                assert abs(got - ref) <= 1e-10 * abs(ref), (name, got, ref)
        seen.add(_ll_class(ref))
    assert seen == {"nan", "-inf", "finite"}
    c.close()


def _same(a, b):
    return (np.isnan(a) and np.isnan(b)) or a == b


@pytest.mark.parametrize("K", [3, 8])
def test_loglikelihood2_equals_two_single_passes(oa, O, K):
    """admx_loglikelihood2 (both states' log-likelihoods in one pass over the
    genotypes, as admixture_qn! pairs ll_basic and ll_qn) returns exactly what
    two admx_loglikelihood calls return, bit for bit, for every pairing of
    plain and degenerate states: both clear (the fused product-form kernel),
    one flagged (the other state falls back to its single-state product form),
    both flagged (term by term)."""
    note("loglikelihood2 pairs", K)
    I, J = 300, 64
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.02, seed=61)
    B = bed.reshape((pitch, J), order="F").copy()
    B[: I // 4, 21] = 0x00  # SNP 21: monomorphic, every genotype 0
    bed2 = np.asfortranarray(B).ravel(order="F")
    c = ctx_for(oa, bed2, pitch, I, J, K)
    q, p = random_state(I, J, K, 7)
    q2, p2 = random_state(I, J, K, 8)
    states = {"plain": (q, p), "plain2": (q2, p2)}
    for name, (cols, val) in {"zero_p_mono_snp": ([21], 0.0), "one_p": ([7], 1.0), "nan_p": ([9], np.nan)}.items():
        pp = p2.copy()
        pp[:, cols] = val
        states[name] = (q2, np.asfortranarray(pp))
    for n1, (qa, pa) in states.items():
        for n2, (qb, pb) in states.items():
            c.set_x(2, np.asfortranarray(np.hstack([qa, pa])))
            c.set_x(3, np.asfortranarray(np.hstack([qb, pb])))
            r1, r2 = c.loglikelihood(2), c.loglikelihood(3)
            f1, f2 = c.loglikelihood2(2, 3)
            print(n1, n2, r1, r2, f1, f2, flush=True)
            assert _same(f1, r1) and _same(f2, r2), (n1, n2)
    c.close()


def test_loglikelihood2_bf16_form_takes_one_pass_per_state(oa, O):
    """In the BF16 form each state's log-likelihood runs on the matrix cores in
    its own pass: loglikelihood2 returns exactly the two single calls."""
    note("loglikelihood2 bf16")
    I, J, K = 1500, 4100, 8
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.02, seed=17)
    c = ctx_for(oa, bed, pitch, I, J, K)
    c.set_precision("bf16")
    q, p = random_state(I, J, K, 5)
    q2, p2 = random_state(I, J, K, 6)
    c.set_x(0, np.asfortranarray(np.hstack([q, p])))
    c.set_x(1, np.asfortranarray(np.hstack([q2, p2])))
    f1, f2 = c.loglikelihood2(0, 1)
    assert f1 == c.loglikelihood(0) and f2 == c.loglikelihood(1)
    c.close()


@pytest.mark.parametrize("I,J,K,T", [(1003, 2021, 8, F64), (4099, 257, 5, F64), (300, 40, 12, F64), (777, 1500, 6, F32),
                                     (13, 7, 3, F64), (1, 1, 2, F64)])
def test_loglikelihood2_random_states(oa, O, I, J, K, T):
    """Two-state log-likelihood on ragged cohorts (fused up to K = 8): equal to
    the two single passes bit for bit and to the oracle within 1e-10 (Float64)
    or 1e-5 (a Float32 context, whose passes run on FP64 copies of its state)."""
    note("loglikelihood2", I, J, K, T.__name__)
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.03, seed=I + K)
    c = ctx_for(oa, bed, pitch, I, J, K, T=T)
    q, p = random_state(I, J, K, 3)
    q2, p2 = random_state(I, J, K, 4)
    c.set_x(0, np.asfortranarray(np.hstack([q, p]), dtype=T))
    c.set_x(1, np.asfortranarray(np.hstack([q2, p2]), dtype=T))
    f1, f2 = c.loglikelihood2(0, 1)
    assert f1 == c.loglikelihood(0) and f2 == c.loglikelihood(1)
    tol = 1e-10 if T is F64 else 1e-5
    for got, (qq, pp) in ((f1, (q, p)), (f2, (q2, p2))):
        ref = O.loglikelihood(bed, pitch, np.asfortranarray(qq.astype(T).astype(F64)),
                              np.asfortranarray(pp.astype(T).astype(F64)))
        assert abs(got - ref) <= tol * abs(ref)
    c.close()


def test_em_on_a_monomorphic_snp_matches_reference(oa, O):
    """FRAPPE EM from a random start on a cohort with a SNP whose genotypes
    are all 0: p of that SNP becomes exactly 0 after the first step, so the
    reference's log-likelihood is NaN from there on; the device's EM
    log-likelihoods match the oracle's, NaN for NaN."""
    note("EM monomorphic")
    I, J, K = 200, 300, 3
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.0, seed=62)
    B = bed.reshape((pitch, J), order="F").copy()
    B[: I // 4, 40] = 0x00
    bed2 = np.asfortranarray(B).ravel(order="F")
    d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(5))
    c = ctx_for(oa, bed2, pitch, I, J, K)
    ref = O.OracleAdmix(bed2, pitch, I, J, K, 3, d.x, d.U, d.V)
    lls = oa.init_em(d, None, 3, d_cu=c, log=None)
    rlls = ref.init_em(3)
    print(lls, rlls, flush=True)
    assert [_ll_class(v) for v in lls] == [_ll_class(v) for v in rlls]
    assert _ll_class(rlls[0]) == "nan"
    c.close()


@pytest.mark.parametrize("K", [9, 12])
def test_hessian_split_switch_is_exact(oa, O, K):
    """ADMX_TUNE_SPLIT_Q / _SPLIT_P: the FP64 update passes with the Hessian
    triangle in one launch or in two (each launch owning half of the
    accumulators) give the same XtX, Xtz, QP steps and projected iterates bit
    for bit, and match the oracle."""
    from openadmixture_jl_amd import _native as N

    note("hessian split", K)
    I, J = 500, 1500
    d, c, ref = _pair(oa, O, I, J, K, seed=K + 40)
    ref.init_em(2)
    x = np.asfortranarray(ref.x.copy())
    out = {}
    for split in (0, 1):
        c.set_tuning(N.ADMX_TUNE_SPLIT_Q, split)
        c.set_tuning(N.ADMX_TUNE_SPLIT_P, split)
        c.set_x(0, x)
        c.update_q(False)
        hq = c.hessians(0)
        c.set_x(1, x)
        c.update_p(False)
        hp = c.hessians(1)
        out[split] = (hq, hp, c.get_x(1))
    for a, b in zip(out[0][0] + out[0][1], out[1][0] + out[1][1]):
        assert np.array_equal(a, b)
    assert np.array_equal(out[0][2], out[1][2])
    ref.update_q(False)
    ok, e = _colmax_close(out[1][0][0], ref.XtX_q, 1e-9)
    assert ok, e
    # the gradient formed from the triangle, sum b p and sum d after the pass,
    # at an EM state whose q columns do not sum to 1
    ok, e = _colmax_close(out[1][0][1], ref.Xtz_q, 1e-9)
    assert ok, e
    c.close()


def test_autotune_sets_waves_and_split(oa, O):
    """autotune() picks a waves candidate and, at K >= 9, a Hessian split per
    side, sets them on the context and caches them per shape."""
    from openadmixture_jl_amd import _native as N

    note("autotune")
    I, J, K = 700, 1600, 10
    d, c, ref = _pair(oa, O, I, J, K, seed=3)
    c.set_x(0, d.x)
    t = oa.autotune(c, candidates=(8192, 65536), reps=1)
    assert t["waves"] in (8192, 65536) and t["split_q"] in (0, 1) and t["split_p"] in (0, 1)
    assert t["mfma_q"] == -1 and t["mfma_p"] == -1
    assert oa.autotune(c, candidates=(8192, 65536), reps=1) == t
    # the BF16 form tunes which kernel runs each side instead of the split
    c.set_precision("bf16")
    tb = oa.autotune(c, candidates=(8192, 65536), reps=1)
    assert tb["mfma_q"] in (0, 1) and tb["mfma_p"] in (0, 1) and tb["split_q"] == -1
    assert oa.autotune(c, candidates=(8192, 65536), reps=1) == tb
    c.close()
