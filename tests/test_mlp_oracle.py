"""The learned warm start on the CPU: the numpy oracle of the network (pinned
by central finite differences and Adam's closed-form first step), the feature
restatement (pinned by a direct loop), the component matching, and the
judge/fallback decision of warm_started_fit (driven by a scripted context).
No reference counterpart exists for any of these."""
import numpy as np
import pytest

import mlp_oracle as M


def _net(in_dim=5, hid=7, out=4, seed=0):
    rng = np.random.default_rng(seed)
    n = hid * in_dim + hid + hid * hid + hid + out * hid + out
    w = rng.normal(0, 0.5, n)
    X = rng.normal(0, 1, (9, in_dim))
    T = rng.dirichlet(np.ones(out), size=9)
    return w, X, T


def test_oracle_gradient_matches_finite_differences():
    in_dim, hid, out = 5, 7, 4
    w, X, T = _net(in_dim, hid, out)
    L, g = M.loss_grad(w, X, T, in_dim, hid, out)
    assert abs(L - M.loss(w, X, T, in_dim, hid, out)) < 1e-14
    h = 1e-6
    fd = np.empty_like(w)
    for e in range(w.size):
        wp, wm = w.copy(), w.copy()
        wp[e] += h
        wm[e] -= h
        fd[e] = (M.loss(wp, X, T, in_dim, hid, out) - M.loss(wm, X, T, in_dim, hid, out)) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-7 * max(1.0, np.max(np.abs(g)))
    # a dropped term (the softmax-target difference without 1/B) would fail this
    assert np.max(np.abs(fd - g * X.shape[0])) > 1e-3


def test_softmax_cross_entropy_closed_form():
    # zero network: uniform output, loss = log(out) for any target on the simplex
    in_dim, hid, out = 3, 4, 5
    n = hid * in_dim + hid + hid * hid + hid + out * hid + out
    T = np.random.default_rng(1).dirichlet(np.ones(out), size=6)
    L = M.loss(np.zeros(n), np.ones((6, in_dim)), T, in_dim, hid, out)
    assert abs(L - np.log(out)) < 1e-14


def test_adam_first_step_is_lr_times_sign():
    g = np.array([3.0, -0.2, 1e-3])
    w, m, v = M.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=0.01, eps=0.0)
    assert np.allclose(w, -0.01 * np.sign(g), rtol=0, atol=1e-15)


def test_features_restatement_direct_loop():
    rng = np.random.default_rng(3)
    I, J, Jp = 7, 11, 9
    g = rng.integers(-1, 3, size=(I, J)).astype(np.int8)
    R = rng.normal(size=(Jp, 64)).astype(np.float32)
    f = rng.uniform(0.1, 0.9, Jp)
    mu, isd = 2 * f, 1 / np.sqrt(2 * f * (1 - f))
    F = M.features(g, R, mu, isd)
    ref = np.zeros((I, 64))
    for i in range(I):
        for j in range(Jp):
            if g[i, j] >= 0:
                ref[i] += (g[i, j] - mu[j]) * isd[j] * R[j]
    assert np.allclose(F, ref / np.sqrt(Jp), rtol=1e-12, atol=1e-12)


def test_match_components_recovers_permutation(oa):
    from openadmixture_jl_amd.warmstart import match_components, world_frequencies

    P, _ = world_frequencies(6, 3000, fst=0.05, seed=2)
    perm_true = np.array([3, 0, 5, 1])
    noisy = np.clip(P[perm_true] + np.random.default_rng(1).normal(0, 0.02, (4, 3000)), 0.01, 0.99)
    perm, corr = match_components(noisy, P)
    assert np.array_equal(perm, perm_true) and corr.min() > 0.5


def _scripted(oa, ll_warm, ll_em, ll_hw):
    """Stand-in device context that records the decision path of warm_started_fit."""
    from openadmixture_jl_amd.hip import HIPAdmixData

    class Fake(HIPAdmixData):
        def __init__(self):
            self.I, self.J, self.K, self.Q, self.T = 4, 6, 2, 3, np.dtype(np.float64)
            self._h = None
            self.xs = {w: np.full((2, 10), 0.5, order="F") for w in range(4)}  # x, x_next, x_next2, x_tmp
            self.ll = (float("nan"), float("nan"))
            self.calls = []

        @property
        def x(self):
            return self.xs[0]

        def close(self):
            pass

        def synchronize(self):
            pass

        def set_skipmissing(self, skip=True):
            pass

        def features(self, R, mu, isd, out=None):
            self.calls.append("features")
            return out

        def warm_q(self, net):
            self.calls.append("warm_q")
            self.xs[0][:, :4] = 0.9  # marks the proposal

        def warm_p(self):
            self.calls.append("warm_p")
            self.xs[0][:, 4:] = 0.375
            return ll_hw  # the log-likelihood at p_k = f for every k

        def copy_x(self, dst, src):
            self.xs[dst] = self.xs[src].copy(order="F")

        def copy_block(self, dst, src, side):
            cols = slice(0, 4) if side == 0 else slice(4, 10)
            self.xs[dst][:, cols] = self.xs[src][:, cols]

        def update_p(self, update2=False, status=False):
            # the MM step reads q_next, which the warm start sets to the proposal
            assert np.all(self.xs[1][:, :4] == 0.9)
            self.calls.append("update_p")
            self.xs[1][:, 4:] = self.xs[0][:, 4:] + 0.0625

        def get_x(self, which, out=None):
            x = self.xs[which]
            return x.copy(order="F") if out is None else np.copyto(out, x) or out

        def set_x(self, which, x):
            self.xs[which] = np.array(x, order="F")

        def set_uv(self, U, V):
            pass

        def set_ll(self, a, b):
            self.ll = (a, b)

        def get_ll(self):
            return self.ll

        def update_q(self, update2=False, status=False):
            # the fixed-P step of an accepted proposal reads the restored proposal
            assert np.all(self.xs[0][:, :4] == 0.9) and np.all(self.xs[0][:, 4:] == 0.5)
            self.calls.append("update_q")
            self.xs[1][:, :4] = 0.95

        def loglikelihood(self, which=0):
            # the state after the P refinement, then after the fixed-P step
            return ll_warm + (1.0 if "update_q" in self.calls else 0.0)

        def qn_iteration(self, it, fix_q=False, fix_p=False):
            assert not fix_q and not fix_p
            self.calls.append("qn")
            return dict(prev=-1.0, basic=-1.0, qn=-1.0, new=-1.0, reldiff=0.0)

        def init_em(self, iters):
            self.calls.append(f"em{iters}")
            self.ll = (self.ll[1], ll_em)
            self.xs[0][:, :4] = 0.1  # marks the EM path
            return np.full(iters, ll_em)

    return Fake()


class _Net:
    out_dim = 2


# (K - 1) J = 6 for the scripted cohort: the proposal must gain more than 6
# over the Hardy-Weinberg log-likelihood and beat one EM step
@pytest.mark.parametrize("ll_warm,ll_em,ll_hw,chosen", [(-10.0, -20.0, -100.0, "warm"), (-30.0, -20.0, -100.0, "em"),
                                                        (-10.0, -20.0, -15.0, "em")])
def test_judge_keeps_the_better_start(oa, ll_warm, ll_em, ll_hw, chosen):
    from openadmixture_jl_amd.warmstart import Panel, warm_started_fit

    ctx = _scripted(oa, ll_warm, ll_em, ll_hw)
    panel = Panel(np.full(6, 0.3), seed=0)
    r = warm_started_fit(ctx, _Net(), panel, seed=1, max_iter=3, p_steps=2)
    assert r["chosen"] == chosen
    assert r["ll_hw"] == ll_hw and r["gain"] == (ll_warm - ll_hw) / 6
    if chosen == "warm":
        # the proposal is restored after the judge's EM step, then takes one fixed-P step
        assert ctx.calls == ["features", "warm_q", "warm_p", "update_p", "update_p", "em1", "update_q", "qn"]
        assert np.all(ctx.x[:, :4] == 0.95)  # q after the fixed-P step
        assert np.all(ctx.x[:, 4:] == 0.5)  # P after two MM steps
        assert r["ll_warm"] == ll_warm  # the judge saw the P-refined proposal
        assert r["history"][0][1] == ll_warm + 1.0  # the fit starts after the fixed-P step
    else:
        # fallback: EM(5) from the random start
        assert ctx.calls == ["features", "warm_q", "warm_p", "update_p", "update_p", "em1", "em4", "qn"]
        assert np.all(ctx.x[:, :4] == 0.1)
        assert r["history"][0][1] == ll_em
