"""Pins of the CPU oracle (oracle/admix_oracle.c) against what fixes each
function independently of the oracle's own formulas:

  decode/counts   PLINK-1 known-answer bytes and numpy bit unpacking
  LL              math.fsum of the closed form on tiny inputs
  update_p        exact first/second derivatives of LL (finite differences)
  update_q        derivatives of LL plus the ADMIXTURE-form terms
                  Xtz_q = dLL/dq + (sum_j m d) 1,
                  XtX_q = -d2LL/dq2 + sum_j m b (11' - 1p' - p1')
  em_q/em_p       sum_k q_next = n_obs/J (the 2J divisor); EM ascent on
                  missing-free data; p_next in [0,1]
  sweep           sweeping every pivot of an SPD matrix gives -inv(A);
                  inverse sweeps restore A
  QP              exhaustive active-set (KKT) enumeration
  projections     Duchi et al. (2008) sort-and-threshold simplex projection
  ZAL             dense linear algebra (numpy.linalg.solve)
  outer loop      parameter recovery on a synthetic cohort
"""
import itertools
import math

import numpy as np
import pytest

from conftest import make_cohort, random_state


def dense_g(O, bed, pitch, I, J):
    g = O.decode(bed, pitch, I, J).astype(np.float64)
    m = (g >= 0).astype(np.float64)
    g[g < 0] = 0
    return g, m


def ll_np(g, m, q, p):
    s = q.T @ p
    return math.fsum((m * (g * np.log(s) + (2 - g) * np.log(1 - s))).ravel())


# ------------------------------------------------------------------ decode / counts
def test_decode_known_answer(O):
    bed = np.array([0xE4, 0x1B, 0x00, 0x55], dtype=np.uint8)  # 4 SNPs x 4 individuals
    g = O.decode(bed, 1, 4, 4)
    assert g[:, 0].tolist() == [0, -1, 1, 2]
    assert g[:, 1].tolist() == [2, 1, -1, 0]
    assert g[:, 2].tolist() == [0, 0, 0, 0]
    assert g[:, 3].tolist() == [-1, -1, -1, -1]


@pytest.mark.parametrize("I,J", [(1, 1), (5, 3), (37, 53), (1001, 77)])
def test_decode_counts_vs_unpackbits(O, I, J):
    rng = np.random.default_rng(I)
    pitch = (I + 3) // 4
    bed = rng.integers(0, 256, size=pitch * J, dtype=np.uint8)
    bits = np.unpackbits(bed.reshape(J, pitch), axis=1, bitorder="little").reshape(J, pitch * 4, 2)
    codes = (bits[:, :, 0] + 2 * bits[:, :, 1])[:, :I].T  # I x J
    ref = np.choose(codes, [0, -1, 1, 2])
    g = O.decode(bed, pitch, I, J)
    assert np.array_equal(g, ref)
    snm, salt, inm = O.counts(bed, pitch, I, J)
    assert np.array_equal(snm, (ref >= 0).sum(0))
    assert np.array_equal(salt, np.where(ref > 0, ref, 0).sum(0))
    assert np.array_equal(inm, (ref >= 0).sum(1))


def test_generator_statistics(O):
    I, J, K = 400, 3000, 3
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, miss=0.02, seed=3)
    g = O.decode(bed, pitch, I, J).astype(float)
    s = Qt.T @ Pt
    obs = g >= 0
    assert abs((~obs).mean() - 0.02) < 0.002
    # E[g] = 2 s, Var[g] = 2 s (1-s): standardised mean residual ~ N(0, 1/n)
    z = ((g - 2 * s) / np.sqrt(2 * s * (1 - s)))[obs]
    assert abs(z.mean()) < 5 / np.sqrt(z.size)
    assert abs(z.var() - 1) < 0.02
    assert np.array_equal(bed, O.gen_bed(Qt, Pt, 3, 0.02, pitch))


# ------------------------------------------------------------------ log-likelihood
@pytest.mark.parametrize("T,rtol", [(np.float64, 1e-13), (np.float32, 2e-6)])
def test_loglikelihood_closed_form(O, T, rtol):
    I, J, K = 131, 157, 3
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, miss=0.05, seed=8)
    q, p = random_state(I, J, K, 2, T)
    g, m = dense_g(O, bed, pitch, I, J)
    ref = ll_np(g, m, q.astype(np.float64), p.astype(np.float64))
    got = O.loglikelihood(bed, pitch, q, p)
    assert abs(got - ref) <= rtol * abs(ref)
    # sub-ranges add up
    a = O.loglikelihood(bed, pitch, q, p, (0, 70), (0, J)) + O.loglikelihood(bed, pitch, q, p, (70, I), (0, J))
    assert abs(a - got) <= 1e-12 * abs(got)


# ------------------------------------------------------------------ derivatives (update_p / update_q)
def _ll_of(O, bed, pitch, q, p):
    return O.loglikelihood(bed, pitch, np.asfortranarray(q), np.asfortranarray(p), nthreads=1)


def test_update_p_is_exact_gradient_and_hessian(O):
    I, J, K = 60, 5, 3
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, miss=0.05, seed=13)
    q, p = random_state(I, J, K, 4)
    XtX = np.zeros((K, K, J), order="F")
    Xtz = np.zeros((K, J), order="F")
    O.update_p_loop(XtX, Xtz, bed, pitch, q, p)
    h = 1e-5
    for j in range(J):
        for k in range(K):
            pp, pm = p.copy(order="F"), p.copy(order="F")
            pp[k, j] += h
            pm[k, j] -= h
            fd = (_ll_of(O, bed, pitch, q, pp) - _ll_of(O, bed, pitch, q, pm)) / (2 * h)
            assert abs(Xtz[k, j] - fd) <= 1e-6 * (1 + abs(fd))
            # Hessian column k by differencing the analytic gradient
            gp = np.zeros((K, K, J), order="F"), np.zeros((K, J), order="F")
            gm = np.zeros((K, K, J), order="F"), np.zeros((K, J), order="F")
            O.update_p_loop(gp[0], gp[1], bed, pitch, q, pp)
            O.update_p_loop(gm[0], gm[1], bed, pitch, q, pm)
            hcol = (gp[1][:, j] - gm[1][:, j]) / (2 * h)
            assert np.allclose(-hcol, XtX[:, k, j], rtol=1e-5, atol=1e-6 * np.abs(XtX[:, :, j]).max())


def test_update_q_admixture_form(O):
    I, J, K = 4, 300, 3
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, miss=0.05, seed=17)
    q, p = random_state(I, J, K, 6)
    XtX = np.zeros((K, K, I), order="F")
    Xtz = np.zeros((K, I), order="F")
    O.update_q_loop(XtX, Xtz, bed, pitch, q, p)
    g, m = dense_g(O, bed, pitch, I, J)
    h = 1e-6
    for i in range(I):
        s = q[:, i] @ p
        d = m[i] * (2 - g[i]) / (1 - s)
        b = m[i] * (2 - g[i]) / (1 - s) ** 2
        grad = np.zeros(K)
        for k in range(K):
            qp_, qm_ = q.copy(order="F"), q.copy(order="F")
            qp_[k, i] += h
            qm_[k, i] -= h
            grad[k] = (_ll_of(O, bed, pitch, qp_, p) - _ll_of(O, bed, pitch, qm_, p)) / (2 * h)
        assert np.allclose(Xtz[:, i], grad + d.sum(), rtol=1e-6, atol=1e-4)
        # second derivative by differencing the gradient identity
        H = np.zeros((K, K))
        for k in range(K):
            qp_, qm_ = q.copy(order="F"), q.copy(order="F")
            qp_[k, i] += h
            qm_[k, i] -= h
            zp = np.zeros((K, K, I), order="F"), np.zeros((K, I), order="F")
            zm = np.zeros((K, K, I), order="F"), np.zeros((K, I), order="F")
            O.update_q_loop(zp[0], zp[1], bed, pitch, qp_, p)
            O.update_q_loop(zm[0], zm[1], bed, pitch, qm_, p)
            # d/dq_k of (Xtz - sum d) = d2LL/dq dq_k
            sp = qp_[:, i] @ p
            sm = qm_[:, i] @ p
            dp = (m[i] * (2 - g[i]) / (1 - sp)).sum()
            dm = (m[i] * (2 - g[i]) / (1 - sm)).sum()
            H[:, k] = ((zp[1][:, i] - dp) - (zm[1][:, i] - dm)) / (2 * h)
        corr = np.zeros((K, K))
        for j in range(J):
            pj = p[:, j]
            corr += b[j] * (np.ones((K, K)) - np.outer(np.ones(K), pj) - np.outer(pj, np.ones(K)))
        assert np.allclose(XtX[:, :, i], -H + corr, rtol=1e-5, atol=1e-5 * np.abs(XtX[:, :, i]).max())


# ------------------------------------------------------------------ EM
def test_em_q_row_sums_and_em_p_range(O):
    I, J, K = 90, 400, 3
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, miss=0.1, seed=23)
    q, p = random_state(I, J, K, 8)
    x0 = np.asfortranarray(np.hstack([q, p]))
    st = O.OracleAdmix(bed, pitch, I, J, K, 3, x0)
    st.em_q()
    st.em_p()
    q_next = st.x_next[:, :I]
    _, m = dense_g(O, bed, pitch, I, J)
    assert np.allclose(q_next.sum(0), m.sum(1) / J, rtol=1e-12)
    p_next = st.x_next[:, I:]
    assert np.all((p_next >= 0) & (p_next <= 1))


def test_em_ascent_without_missing(O):
    I, J, K = 120, 600, 3
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, miss=0.0, seed=29)
    q, p = random_state(I, J, K, 9)
    st = O.OracleAdmix(bed, pitch, I, J, K, 3, np.hstack([q, p]))
    prev = st.loglikelihood()
    lls = st.init_em(6)
    assert lls[0] > prev
    assert np.all(np.diff(lls) > 0)


# ------------------------------------------------------------------ sweep / QP
def test_sweep_gives_negative_inverse(O):
    rng = np.random.default_rng(1)
    A0 = rng.normal(size=(6, 6))
    A0 = A0 @ A0.T + 6 * np.eye(6)
    A = np.asfortranarray(A0.copy())
    for k in range(6):
        O.sweep(A, k)
    assert np.allclose(A, -np.linalg.inv(A0), rtol=1e-10, atol=1e-12)
    for k in range(6):
        O.sweep(A, k, inverse=True)
    assert np.allclose(A, A0, rtol=1e-10, atol=1e-10)


def _qp_bruteforce(H, gvec, x, simplex):
    """min 0.5 d'Hd - gvec'd  s.t. 0 <= x+d <= 1 (and 1'd = 1 - 1'x): enumerate
    the active bound sets, solve each equality-constrained KKT system, keep the
    best feasible point."""
    K = len(x)
    best, bestv = None, np.inf
    for pattern in itertools.product((0, 1, 2), repeat=K):  # 0 free, 1 at lower, 2 at upper
        fixed = {k: (-x[k] if s == 1 else 1 - x[k]) for k, s in enumerate(pattern) if s}
        free = [k for k in range(K) if not pattern[k]]
        d = np.zeros(K)
        for k, v in fixed.items():
            d[k] = v
        if free:
            Hf = H[np.ix_(free, free)]
            rhs = gvec[free] - H[np.ix_(free, list(fixed))] @ np.array(list(fixed.values())) if fixed else gvec[free]
            if simplex:
                n = len(free)
                Kk = np.zeros((n + 1, n + 1))
                Kk[:n, :n] = Hf
                Kk[:n, n] = 1
                Kk[n, :n] = 1
                rr = np.append(rhs, 1 - x.sum() - sum(fixed.values()))
                try:
                    sol = np.linalg.solve(Kk, rr)
                except np.linalg.LinAlgError:
                    continue
                d[free] = sol[:n]
            else:
                try:
                    d[free] = np.linalg.solve(Hf, rhs)
                except np.linalg.LinAlgError:
                    continue
        elif simplex and abs(d.sum() - (1 - x.sum())) > 1e-12:
            continue
        xn = x + d
        if np.any(xn < -1e-12) or np.any(xn > 1 + 1e-12):
            continue
        if simplex and abs(xn.sum() - 1) > 1e-9:
            continue
        v = 0.5 * d @ H @ d - gvec @ d
        if v < bestv:
            best, bestv = d, v
    return best, bestv


@pytest.mark.parametrize("simplex", [True, False])
@pytest.mark.parametrize("K", [2, 3, 4, 5])
def test_qp_matches_kkt_enumeration(O, simplex, K):
    I, J = 40, 60
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, miss=0.02, seed=31 + K)
    q, p = random_state(I, J, K, K)
    if simplex:
        XtX, Xtz = np.zeros((K, K, I), order="F"), np.zeros((K, I), order="F")
        O.update_q_loop(XtX, Xtz, bed, pitch, q, p)
        xs = q
    else:
        XtX, Xtz = np.zeros((K, K, J), order="F"), np.zeros((K, J), order="F")
        O.update_p_loop(XtX, Xtz, bed, pitch, q, p)
        xs = p
    worst = 0.0
    for col in range(xs.shape[1]):
        H, gv, x = XtX[:, :, col].copy(), Xtz[:, col].copy(), xs[:, col].copy()
        tab = O.create_tableau(np.asfortranarray(H), -gv, x, simplex)
        d, rc = O.quadratic_program(tab, x, np.zeros(K), np.ones(K), 1 if simplex else 0)
        bf, bv = _qp_bruteforce(H, gv, x, simplex)
        if rc == 0:
            continue
        v = 0.5 * d @ H @ d - gv @ d
        # active-set solution: feasible, and optimal up to the method's 1e-5 bound tolerance
        assert np.all(x + d >= -1e-12) and np.all(x + d <= 1 + 1e-12)
        if simplex:
            assert abs((x + d).sum() - 1) < 1e-9
        scale = 1e-5 * (np.abs(gv).max() + np.abs(H).max())
        assert v <= bv + scale, (col, v, bv)
        worst = max(worst, np.abs(d - bf).max())
    assert worst < 1e-3


def test_qp_failed_pivot_returns_zero(O):
    K = 3
    H = np.asfortranarray(np.diag([1.0, -1.0, 2.0]))
    tab = O.create_tableau(H, np.array([0.1, 0.2, 0.3]), np.array([0.3, 0.3, 0.4]), True)
    d, rc = O.quadratic_program(tab, np.array([0.3, 0.3, 0.4]), np.zeros(K), np.ones(K), 1)
    assert rc == 0 and np.all(d == 0)


# ------------------------------------------------------------------ projections
def _duchi(v, z):
    """Euclidean projection of v onto {w >= 0, sum w = z} (Duchi et al., ICML 2008)."""
    u = np.sort(v)[::-1]
    css = np.cumsum(u)
    rho = np.nonzero(u * np.arange(1, len(v) + 1) > (css - z))[0][-1]
    theta = (css[rho] - z) / (rho + 1.0)
    return np.maximum(v - theta, 0)


@pytest.mark.parametrize("K", [1, 2, 3, 8, 16])
def test_project_q_is_simplex_projection(O, K):
    rng = np.random.default_rng(K)
    b = np.asfortranarray(rng.normal(0.3, 0.7, size=(K, 500)))
    out = O.project_q(b.copy(order="F"))
    pc = 1e-5
    for col in range(b.shape[1]):
        ref = _duchi(b[:, col] - pc, 1 - K * pc) + pc
        assert np.allclose(out[:, col], ref, atol=1e-13)
    assert np.allclose(out.sum(0), 1, atol=1e-12)
    assert np.all(out >= pc - 1e-15)
    again = O.project_q(out.copy(order="F"))
    assert np.allclose(again, out, atol=1e-14)


def test_project_p_clamps(O):
    v = np.array([-1.0, 0.0, 1e-6, 0.5, 1 - 1e-6, 1.0, 3.0])
    out = O.project_p(v.copy())
    assert np.array_equal(out, np.clip(v, 1e-5, 1 - 1e-5))


# ------------------------------------------------------------------ ZAL quasi-Newton
@pytest.mark.parametrize("m", [1, 2, 3])
def test_update_qn_dense_solve(O, m):
    rng = np.random.default_rng(m)
    N = 500
    x, xm = rng.random(N), rng.random(N)
    U = np.asfortranarray(rng.normal(size=(N, m)))
    V = np.asfortranarray(rng.normal(size=(N, m)))
    out = np.zeros(N)
    O.update_QN(out, xm, x, U, V, m)
    ref = xm - V @ np.linalg.solve(U.T @ (U - V), U.T @ (x - xm))
    assert np.allclose(out, ref, rtol=1e-10, atol=1e-12)


def test_update_uv_columns(O):
    rng = np.random.default_rng(0)
    N, Q = 50, 3
    x, xn, xn2 = rng.random(N), rng.random(N), rng.random(N)
    U = np.asfortranarray(np.zeros((N, Q)))
    V = np.asfortranarray(np.zeros((N, Q)))
    O.update_UV(U, V, x, xn, xn2, 5, Q)  # column (5-1) % 3 = 1
    assert np.array_equal(U[:, 1], xn - x) and np.array_equal(V[:, 1], xn2 - xn)
    assert not U[:, 0].any() and not U[:, 2].any()


# ------------------------------------------------------------------ outer loop
def test_outer_loop_recovers_truth(O):
    I, J, K = 300, 3000, 3
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, alpha=0.5, miss=0.01, seed=41)
    rng = np.random.default_rng(2)
    x0 = rng.random((K, I + J))
    x0[:, :I] /= x0[:, :I].sum(0)
    st = O.OracleAdmix(bed, pitch, I, J, K, 3, x0, rng.random((K * (I + J), 3)), rng.random((K * (I + J), 3)))
    st.init_em(5)
    last = None
    for it in range(1, 60):
        r = st.qn_iteration(it)
        last = r
        if r["reldiff"] < 1e-7:
            break
    assert last["reldiff"] < 1e-7
    q = st.x[:, :I]
    err = min(np.abs(q[list(pm)] - Qt).mean() for pm in itertools.permutations(range(K)))
    assert err < 0.03
    truth_ll = O.loglikelihood(bed, pitch, np.asfortranarray(Qt), np.asfortranarray(Pt))
    assert st.ll_new > truth_ll  # the MLE beats the generating parameters


def test_f32_oracle_tracks_f64(O):
    I, J, K = 200, 800, 3
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, seed=43)
    q, p = random_state(I, J, K, 1)
    a = O.loglikelihood(bed, pitch, q, p)
    b = O.loglikelihood(bed, pitch, q.astype(np.float32), p.astype(np.float32))
    assert abs(a - b) <= 1e-5 * abs(a)


# ---------------------------------------------------------------- LBQN (src/quasi_newton.jl:15-58)
@pytest.mark.parametrize("m_hist", [1, 3])
def test_lbqn_jumps_to_the_fixed_point_of_a_scalar_linear_map(O, m_hist):
    """For F(x) = a x + b (any dimension), every V column is (a - 1) times its U
    column, so the two-loop recursion gives x - u/(a-1) = b/(1-a) exactly."""
    rng = np.random.default_rng(4)
    N, Q, a = 50, 3, 0.7
    b = rng.normal(size=N)
    U = np.asfortranarray(rng.normal(size=(N, Q)))
    V = np.asfortranarray(rng.normal(size=(N, Q)))
    x = rng.normal(size=N)
    for it in range(1, m_hist + 1):  # fill the history with iterates of the same map
        x1 = a * x + b
        x2 = a * x1 + b
        O.update_UV_LBQN(U, V, x, x1, x2, it, Q)
        if it < m_hist:
            x = x2
    m = min(m_hist, Q)
    out = np.empty(N)
    O.update_QN_LBQN(out, x, U, V, m)
    assert np.allclose(out, b / (1 - a), rtol=0, atol=1e-12)


def test_lbqn_history_shift(O):
    rng = np.random.default_rng(5)
    N, Q = 20, 3
    U = np.asfortranarray(rng.normal(size=(N, Q)))
    V = np.asfortranarray(rng.normal(size=(N, Q)))
    U0, V0 = U.copy(order="F"), V.copy(order="F")
    x, x1, x2 = rng.normal(size=(3, N))
    O.update_UV_LBQN(U, V, x, x1, x2, 5, Q)  # iter >= Q: columns shift right, the last drops
    assert np.array_equal(U[:, 1:], U0[:, :2]) and np.array_equal(V[:, 1:], V0[:, :2])
    assert np.array_equal(U[:, 0], x1 - x)
    assert np.array_equal(V[:, 0], (x2 - x1) - (x1 - x))
    U, V = U0.copy(order="F"), V0.copy(order="F")
    O.update_UV_LBQN(U, V, x, x1, x2, 2, Q)  # iter 2: only column 1 shifts into 2
    assert np.array_equal(U[:, 1], U0[:, 0]) and np.array_equal(U[:, 2], U0[:, 2])


# ---------------------------------------------------------------- mean-imputation loops (loops.jl:493-921)
def test_snp_means_from_decoded_genotypes(O):
    I, J = 203, 77
    _, _, bed, pitch = make_cohort(O, I, J, 3, miss=0.1, seed=31)
    g = O.decode(bed, pitch, I, J).astype(np.float64)
    obs = g >= 0
    ref = np.where(obs, g, 0).sum(0) / obs.sum(0)
    assert np.allclose(O.snp_means(bed, pitch, I, J), ref, rtol=1e-15, atol=0)


def test_imputation_loops_equal_skipmissing_without_missing(O):
    I, J, K = 90, 70, 3
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.0, seed=32)
    q, p = random_state(I, J, K, 2)
    mu = O.snp_means(bed, pitch, I, J)
    assert O.loglikelihood(bed, pitch, q, p, mu=mu) == O.loglikelihood(bed, pitch, q, p)
    a = [np.zeros((K, K, I), order="F"), np.zeros((K, I), order="F")]
    b = [np.zeros((K, K, I), order="F"), np.zeros((K, I), order="F")]
    O.update_q_loop(*a, bed, pitch, q, p, mu=mu)
    O.update_q_loop(*b, bed, pitch, q, p)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_imputation_ll_adds_the_missing_entries_at_their_mean(O):
    I, J, K = 80, 60, 3
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.15, seed=33)
    q, p = random_state(I, J, K, 3)
    mu = O.snp_means(bed, pitch, I, J)
    g = O.decode(bed, pitch, I, J)
    s = q.T @ p
    miss = g < 0
    extra = math.fsum((mu[None, :] * np.log(s) + (2 - mu[None, :]) * np.log(1 - s))[miss])
    got = O.loglikelihood(bed, pitch, q, p, mu=mu, nthreads=1)
    ref = O.loglikelihood(bed, pitch, q, p, nthreads=1) + extra
    assert abs(got - ref) <= 1e-12 * abs(ref)


def test_imputation_update_p_is_exact_gradient(O):
    I, J, K = 60, 4, 3
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.1, seed=34)
    q, p = random_state(I, J, K, 5)
    mu = O.snp_means(bed, pitch, I, J)
    XtX = np.zeros((K, K, J), order="F")
    Xtz = np.zeros((K, J), order="F")
    O.update_p_loop(XtX, Xtz, bed, pitch, q, p, mu=mu)
    h = 1e-5
    for j in range(J):
        for k in range(K):
            pp, pm = p.copy(order="F"), p.copy(order="F")
            pp[k, j] += h
            pm[k, j] -= h
            fd = (O.loglikelihood(bed, pitch, q, pp, mu=mu, nthreads=1)
                  - O.loglikelihood(bed, pitch, q, pm, mu=mu, nthreads=1)) / (2 * h)
            assert abs(Xtz[k, j] - fd) <= 1e-6 * (1 + abs(fd))
