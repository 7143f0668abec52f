"""Learned warm start on the device: features, network forward/backward and
Adam against the CPU oracle (oracle/mlp_oracle.py), the projected proposal,
and what the trained network gains over the method without it."""
import numpy as np
import pytest

import mlp_oracle as M
from conftest import make_cohort

pytestmark = pytest.mark.gpu


def _blocks(net):
    i, h, o = net.in_dim, net.hidden, net.out_dim
    sizes = [h * i, h, h * h, h, o * h, o]
    return np.cumsum([0] + sizes)


def test_mlp_loss_and_gradient_match_oracle(oa):
    print("mlp gradient", flush=True)
    rng = np.random.default_rng(0)
    net = oa.MLP(64, 256, 7, seed=3)
    X = rng.normal(size=(300, 64)).astype(np.float32)
    T = rng.dirichlet(0.5 * np.ones(7), size=300).astype(np.float32)
    loss, g = net.loss_grad(X, T)
    w = net.params().astype(np.float64)
    rl, rg = M.loss_grad(w, X, T, 64, 256, 7)
    assert abs(loss - rl) <= 1e-5 * abs(rl)
    off = _blocks(net)
    for b in range(6):  # per parameter block, relative to the block's largest entry
        d = np.abs(g[off[b]:off[b + 1]] - rg[off[b]:off[b + 1]]).max()
        assert d <= 1e-4 * np.abs(rg[off[b]:off[b + 1]]).max() + 1e-9, (b, d)
    net.close()


def test_mlp_predict_matches_oracle(oa):
    rng = np.random.default_rng(1)
    net = oa.MLP(64, 256, 10, seed=4)
    X = rng.normal(size=(70001, 64)).astype(np.float32)  # two inference batches
    Y = net.predict(X)
    ref = M.forward(net.params().astype(np.float64), X, 64, 256, 10)["Y"]
    assert np.abs(Y - ref).max() < 2e-6
    assert np.allclose(Y.sum(1), 1.0, atol=1e-5)
    net.close()


def test_mlp_adam_steps_match_oracle(oa):
    print("mlp adam", flush=True)
    rng = np.random.default_rng(2)
    net = oa.MLP(64, 32, 5, seed=5)
    X = rng.normal(size=(200, 64)).astype(np.float32)
    T = rng.dirichlet(np.ones(5), size=200).astype(np.float32)
    w = net.params().astype(np.float64)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    hist = net.train(X, T, epochs=3, batch=200, lr=1e-2, seed=9)  # full batch: 3 Adam steps
    for t in range(1, 4):
        _, g = M.loss_grad(w, X, T, 64, 32, 5)
        w, m, v = M.adam_step(w, g, m, v, t, 1e-2)
    got = net.params().astype(np.float64)
    assert np.abs(got - w).max() < 1e-4, np.abs(got - w).max()
    assert hist[2] < hist[0]
    net.close()


@pytest.mark.parametrize("I,J,Jp", [(1000, 3000, 2500), (1001, 2003, 2003), (37, 13, 13), (300, 500, 1)])
def test_features_match_oracle(oa, O, I, J, Jp):
    print("features", I, J, Jp, flush=True)
    Qt, Pt, bed, pitch = make_cohort(O, I, J, 3, miss=0.02, seed=3)
    c = oa.HIPAdmixData(I, J, 3, 3)
    c.upload_bed(bed.reshape((pitch, J), order="F"))
    panel = oa.Panel(Pt[:, :Jp].mean(0), seed=1)
    F = panel.features(c)
    g = O.decode(bed, pitch, I, J)
    ref = M.features(g, panel.R, panel.mu.astype(np.float64), panel.isd.astype(np.float64))
    assert np.abs(F - ref).max() <= 1e-5 * np.abs(ref).max()
    # the proposal: simplex projection (pseudocount 1e-5) of the network output
    net = oa.MLP(64, 64, 3, seed=2)
    c.warm_q(net)
    q = c.get_x(0)[:, :I]
    Y = net.predict(F).T.astype(np.float64)
    Yp = np.asfortranarray(Y.copy())
    O.project_q(Yp)
    assert np.abs(q - Yp).max() < 1e-6
    assert np.allclose(q.sum(0), 1.0) and q.min() >= 1e-5 - 1e-12
    net.close()
    c.close()


def test_warm_start_gains_over_random_init(oa, O):
    """Train on simulated cohorts of a 5-population world, evaluate on a new
    cohort of all 5: the proposal must be kept by the judge and reach the plain
    method's final LL (|dLL| < 1e-4 relative) in fewer iterations than the
    random start + EM(5)."""
    print("warm start gain", flush=True)
    Kw, Jp, Jeval, Ieval = 5, 4000, 8000, 6000
    P, anc = oa.world_frequencies(Kw, Jeval, fst=0.05, seed=11)
    panel = oa.Panel(P[:, :Jp].mean(0), seed=3)
    X, T, info = oa.training_set(P, panel, n_cohorts=8, I=1500, J=Jp, Ks=[2, 3, 4, 5], seed=5)
    net = oa.MLP(64, 256, Kw, seed=1)
    hist = net.train(X, T, epochs=30, batch=256, lr=2e-3, seed=1)
    print("train loss", hist[0], hist[-1], flush=True)
    assert hist[-1] < hist[0]

    rng = np.random.default_rng(21)
    Qt = np.asfortranarray(rng.dirichlet(0.5 * np.ones(Kw), size=Ieval).T)
    res = {}
    for mode in ("plain", "warm"):
        c = oa.HIPAdmixData(Ieval, Jeval, Kw, 3)
        c.generate_bed(Qt, P, 77, 0.01)
        if mode == "plain":
            d = oa.fit_on_device(c, Kw, 5, max_iter=500)
            res[mode] = [(0, d.llhist[0])] + [(n, ll) for n, ll in enumerate(d.llhist[1:], 1)]
        else:
            r = oa.warm_started_fit(c, net, panel, seed=5, max_iter=500)
            assert r["chosen"] == "warm", r
            res[mode] = [(n, ll) for n, (t, ll) in enumerate(r["history"])]
        c.close()
    ll_star = res["plain"][-1][1]

    def first(h):
        return next(n for n, ll in h if abs(ll - ll_star) <= 1e-4 * abs(ll_star))

    n_plain, n_warm = first(res["plain"]), first(res["warm"])
    print(f"iterations to |dLL|<1e-4: plain {n_plain}, warm {n_warm}; final LL plain {ll_star:.6e} "
          f"warm {res['warm'][-1][1]:.6e}", flush=True)
    assert n_warm < n_plain
    assert abs(res["warm"][-1][1] - ll_star) <= 1e-6 * abs(ll_star)
    net.close()


def test_warm_p_and_state_copy(oa, O):
    """admx_warm_p: every p[k, j] is SNP j's observed allele frequency
    alt / max(2 nonmissing, 1), clipped to [1e-5, 1 - 1e-5] (checked against the
    oracle's counts, bit for bit, including an all-missing and a monomorphic
    SNP); q is untouched.  The log-likelihood it returns equals the closed form
    sum alt log f + (2 nm - alt) log(1 - f) on the oracle's counts and a
    genotype pass at the new state (q on the simplex, so s = f), both within
    1e-12.  admx_copy_x copies a state vector exactly."""
    print("warm p", flush=True)
    I, J, K = 333, 700, 4
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.05, seed=71)
    B = bed.reshape((pitch, J), order="F").copy()
    B[: (I + 3) // 4, 10] = 0x55  # all missing
    B[: (I + 3) // 4, 11] = 0xFF  # every genotype 2
    B = np.asfortranarray(B)
    c = oa.HIPAdmixData(I, J, K, 3)
    c.upload_bed(B)
    nm, alt, _ = O.counts(np.asfortranarray(B).ravel(order="F"), pitch, I, J)
    x0 = np.asfortranarray(np.random.default_rng(2).random((K, I + J)))
    x0[:, :I] /= x0[:, :I].sum(0)
    c.set_x(0, x0)
    ll_hw = c.warm_p()
    x = c.get_x(0)
    f = np.clip(alt / np.maximum(2 * nm, 1), 1e-5, 1 - 1e-5)
    assert np.array_equal(x[:, I:], np.tile(f, (K, 1)))
    assert np.array_equal(x[:, :I], x0[:, :I])
    assert f[10] == 1e-5 and f[11] == 1 - 1e-5
    nm64, alt64 = nm.astype(np.float64), alt.astype(np.float64)
    ll_ref = float(np.sum(alt64 * np.log(f) + (2 * nm64 - alt64) * np.log1p(-f)))
    ll_pass = c.loglikelihood(0)
    print(f"warm_p LL {ll_hw:.15e}, closed form {ll_ref:.15e}, pass {ll_pass:.15e}", flush=True)
    assert abs(ll_hw - ll_ref) <= 1e-12 * abs(ll_ref)
    assert abs(ll_hw - ll_pass) <= 1e-12 * abs(ll_pass)
    c.copy_x(3, 0)
    assert np.array_equal(c.get_x(3), x)
    c.close()


@pytest.mark.parametrize("prec", ["f64", "bf16"])
def test_warm_p_refinement_matches_oracle(oa, O, prec):
    """The warm start's P refinement as warm_started_fit drives it: q_next = q
    (Q fixed), then two MM steps update_p! with p = p_next after each, from p at
    the observed allele frequencies, against the oracle's update_p! on the same
    state (including an all-missing and a monomorphic SNP).  Plain FP64: p
    within 1e-9 max-abs and the log-likelihood within 1e-10; BF16 form: 1e-3
    and 1e-6 (the per-step bar of the BF16 form's update passes, whose Hessian
    and gradient carry its error)."""
    print(f"warm P refinement {prec}", flush=True)
    I, J, K = 600, 900, 4
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.03, seed=73)
    B = bed.reshape((pitch, J), order="F").copy()
    B[: (I + 3) // 4, 10] = 0x55  # all missing
    B[: (I + 3) // 4, 11] = 0xFF  # every genotype 2
    B = np.asfortranarray(B)
    flat = B.ravel(order="F")
    c = oa.HIPAdmixData(I, J, K, 3)
    c.upload_bed(B)
    c.set_precision(prec)
    x0 = np.asfortranarray(np.random.default_rng(4).dirichlet(np.ones(K), size=I + J).T)
    c.set_x(0, x0)
    c.warm_p()
    x = c.get_x(0)
    o = O.OracleAdmix(flat, pitch, I, J, K, 3, x)
    c.copy_block(1, 0, 0)
    o.x_next[:, :I] = o.x[:, :I]
    for _ in range(2):
        c.update_p()
        c.copy_block(0, 1, 1)
        o.update_p()
        o.x[:, I:] = o.x_next[:, I:]
    xg = c.get_x(0)
    dp = float(np.max(np.abs(xg[:, I:] - o.x[:, I:])))
    ll_g, ll_o = c.loglikelihood(0), o.loglikelihood()
    print(f"  max|dp| {dp:.2e}, LL {ll_g:.12e} vs {ll_o:.12e}", flush=True)
    assert np.array_equal(xg[:, :I], x[:, :I])
    assert dp <= (1e-9 if prec == "f64" else 1e-3)
    assert abs(ll_g - ll_o) <= (1e-10 if prec == "f64" else 1e-6) * abs(ll_o)
    c.close()


def test_mismatched_network_is_rejected(oa, O):
    """A real network trained on one world's populations, used on a cohort of
    unrelated populations (another Balding-Nichols world): its proposal beats
    one EM step from the random start but explains the genotypes no better than
    chance beyond Hardy-Weinberg (gain below (K-1)J), so the judge falls back
    to the random start + EM, and the fit reaches the plain method's LL.  The
    same network on a cohort of drifted copies of its populations (F = 0.05)
    is kept and saves iterations."""
    print("warm start mismatch", flush=True)
    Kw, Jp, J, Ie = 5, 4000, 8000, 6000
    P, _ = oa.world_frequencies(Kw, J, fst=0.05, seed=11)
    panel = oa.Panel(P[:, :Jp].mean(0), seed=3)
    X, T, _ = oa.training_set(P, panel, n_cohorts=8, I=1500, J=Jp, Ks=[2, 3, 4, 5], seed=5)
    net = oa.MLP(64, 256, Kw, seed=1)
    net.train(X, T, epochs=30, batch=256, lr=2e-3, seed=1)
    rng = np.random.default_rng(21)
    Qt = np.asfortranarray(rng.dirichlet(0.5 * np.ones(Kw), size=Ie).T)
    worlds = {"unrelated": oa.world_frequencies(Kw, J, fst=0.05, seed=99)[0],
              "drifted": oa.drifted_populations(P, 0.05, 8)}
    for name, Pe in worlds.items():
        c = oa.HIPAdmixData(Ie, J, Kw, 3)
        c.generate_bed(Qt, np.asfortranarray(Pe), 77, 0.01)
        plain = oa.fit_on_device(c, Kw, 5, max_iter=500)
        ll_star = plain.llhist[-1]
        r = oa.warm_started_fit(c, net, panel, seed=5, max_iter=500)
        c.close()
        lls = [ll for _, ll in r["history"]]
        n_plain = next(n for n, ll in enumerate(plain.llhist) if abs(ll - ll_star) <= 1e-4 * abs(ll_star))
        n_warm = next(n for n, ll in enumerate(lls) if abs(ll - ll_star) <= 1e-4 * abs(ll_star))
        print(f"{name}: chosen {r['chosen']}, gain {r['gain']:.2f} (K-1)J, LL warm start {r['ll_warm']:.6e} vs one "
              f"EM step {r['ll_em']:.6e}; iterations to 1e-4 {n_warm} vs plain {n_plain}", flush=True)
        assert abs(lls[-1] - ll_star) <= 1e-6 * abs(ll_star)
        if name == "unrelated":
            assert r["chosen"] == "em" and r["ll_warm"] > r["ll_em"] and r["gain"] < 1.0
            assert n_warm == n_plain
        else:
            assert r["chosen"] == "warm" and n_warm < n_plain
    net.close()
