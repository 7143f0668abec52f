"""Learned warm start for admixture_qn! (no reference counterpart: the reference
starts every fit from a random point plus FRAPPE EM, src/algorithms_outer.jl:12-45).

Pieces, all computed by libadmixhip kernels:

* ``Panel`` -- the reference panel the network is tied to: the first ``Jp``
  SNPs of a cohort, their panel allele frequencies f_j and a fixed random
  projection R (Jp x 64).  Per individual the feature vector is
  (1/sqrt(Jp)) sum_j (g_ij - 2 f_j) / sqrt(2 f_j (1 - f_j)) R[j, :] over the
  observed panel genotypes (``admx_features``).
* ``MLP`` -- FP32 network 64 -> 256 -> 256 -> K with ReLU and a softmax head,
  trained by Adam on the soft-target cross entropy between its output and the
  converged Q that EM/SQP finds on simulated cohorts (``admx_mlp_*``).
* ``training_set`` -- simulated cohorts drawn from a world of ``Kw`` panel
  populations (each cohort admixes a random subset of K_c of them with
  Dirichlet(alpha) proportions), each solved by the plain method; the fitted
  components are matched to panel populations by their allele frequencies
  (Hungarian assignment on correlation), so targets use the panel's labels.
* ``warm_started_fit`` -- propose q from the network, set p to the observed
  allele frequencies and refine it with Q held fixed, judge the proposal by
  its log-likelihood against one FRAPPE EM step from the random start, fall
  back to the random start + EM when the proposal is worse, then run
  admixture_qn! to convergence.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import _native as N
from .hip import HIPAdmixData
from .structs import AdmixData

FEAT = 64


class MLP:
    """Device MLP in -> hidden -> hidden -> out (FP32) owned by libadmixhip."""

    def __init__(self, in_dim: int = FEAT, hidden: int = 256, out_dim: int = 10, seed: int = 0):
        self.in_dim, self.hidden, self.out_dim = int(in_dim), int(hidden), int(out_dim)
        self._h = C.c_void_p()
        N.check(N.lib().admx_mlp_create(C.byref(self._h), self.in_dim, self.hidden, self.out_dim, C.c_uint64(seed)))
        n = C.c_int64()
        N.check(N.lib().admx_mlp_num_params(self._h, C.byref(n)))
        self.num_params = n.value

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().admx_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        w = np.empty(self.num_params, np.float32)
        N.check(N.lib().admx_mlp_get_params(self._h, N.ptr(w)))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters")
        N.check(N.lib().admx_mlp_set_params(self._h, N.ptr(w)))

    def _rows(self, X, width):
        X = np.ascontiguousarray(X, dtype=np.float32)
        if X.ndim != 2 or X.shape[1] != width:
            raise ValueError(f"expected an (n, {width}) array")
        return X

    def loss_grad(self, X, T):
        X, T = self._rows(X, self.in_dim), self._rows(T, self.out_dim)
        g = np.empty(self.num_params, np.float32)
        loss = C.c_float()
        N.check(N.lib().admx_mlp_loss_grad(self._h, N.ptr(X), N.ptr(T), X.shape[0], C.byref(loss), N.ptr(g)))
        return loss.value, g

    def predict(self, X):
        X = self._rows(X, self.in_dim)
        Y = np.empty((X.shape[0], self.out_dim), np.float32)
        N.check(N.lib().admx_mlp_predict(self._h, N.ptr(X), X.shape[0], N.ptr(Y)))
        return Y

    def train(self, X, T, epochs: int = 20, batch: int = 512, lr: float = 1e-3, seed: int = 0):
        X, T = self._rows(X, self.in_dim), self._rows(T, self.out_dim)
        if X.shape[0] != T.shape[0]:
            raise ValueError("X and T must have the same number of rows")
        hist = np.zeros(max(int(epochs), 1), np.float32)
        N.check(N.lib().admx_mlp_train(self._h, N.ptr(X), N.ptr(T), X.shape[0], int(epochs), int(batch),
                                       C.c_float(lr), C.c_uint64(seed), N.ptr(hist)))
        return hist[: int(epochs)]


class Panel:
    """Reference panel of the features: Jp SNPs, their allele frequencies, R."""

    def __init__(self, freqs, seed: int = 0):
        f = np.clip(np.asarray(freqs, np.float64), 1e-3, 1 - 1e-3)
        self.Jp = f.size
        self.mu = np.ascontiguousarray(2.0 * f, dtype=np.float32)
        self.isd = np.ascontiguousarray(1.0 / np.sqrt(2.0 * f * (1.0 - f)), dtype=np.float32)
        self.R = np.ascontiguousarray(np.random.default_rng(seed).standard_normal((self.Jp, FEAT)), dtype=np.float32)

    def features(self, ctx: HIPAdmixData, to_host: bool = True):
        """Features of ctx's genotypes (kept on the device for warm_q)."""
        F = np.empty((ctx.I, FEAT), np.float32) if to_host else None
        return ctx.features(self.R, self.mu, self.isd, F)


def world_frequencies(Kw: int, J: int, fst: float = 0.05, seed: int = 0):
    """Panel populations by the Balding-Nichols model: ancestral f_j ~ U(0.05, 0.95),
    P[k, j] ~ Beta(f_j (1-F)/F, (1-f_j)(1-F)/F) clipped to [0.01, 0.99]."""
    rng = np.random.default_rng(seed)
    anc = rng.uniform(0.05, 0.95, size=J)
    a = anc * (1 - fst) / fst
    b = (1 - anc) * (1 - fst) / fst
    P = np.clip(rng.beta(a, b, size=(Kw, J)), 0.01, 0.99)
    return np.asfortranarray(P), anc


def drifted_populations(P, fst: float, seed: int = 0):
    """Populations drifted from the rows of P by the Balding-Nichols model:
    P'[k, j] ~ Beta(P[k, j] (1-F)/F, (1 - P[k, j])(1-F)/F), clipped to
    [0.01, 0.99] -- study populations related to, but not the same as, the
    panel populations a network was trained on."""
    rng = np.random.default_rng(seed)
    P = np.asarray(P, np.float64)
    a = P * (1 - fst) / fst
    b = (1 - P) * (1 - fst) / fst
    return np.asfortranarray(np.clip(rng.beta(a, b), 0.01, 0.99))


def match_components(p_fit, p_ref):
    """Permutation perm with fitted component k <-> reference row perm[k],
    maximising the summed correlation of allele frequencies."""
    from scipy.optimize import linear_sum_assignment

    a = p_fit - p_fit.mean(1, keepdims=True)
    b = p_ref - p_ref.mean(1, keepdims=True)
    corr = (a @ b.T) / np.outer(np.linalg.norm(a, axis=1), np.linalg.norm(b, axis=1))
    rows, cols = linear_sum_assignment(-corr)
    perm = np.empty(len(rows), np.int64)
    perm[rows] = cols
    return perm, corr[rows, cols]


def training_set(P_world, panel: Panel, n_cohorts: int, I: int = 5000, J: int = 20000, Ks=range(4, 11),
                 alpha_range=(0.1, 1.0), miss_rate: float = 0.01, seed: int = 0, max_iter: int = 500,
                 log=None):
    """Simulate cohorts from the world, solve each by the plain method and
    return (X, T, info): X the 64 features per individual, T the converged Q in
    the panel's population labels (zero outside the cohort's populations)."""
    from .cv import fit_on_device

    Kw = P_world.shape[0]
    rng = np.random.default_rng(seed)
    Ks = list(Ks)
    X, T, info = [], [], []
    for c in range(n_cohorts):
        Kc = Ks[c % len(Ks)]
        alpha = float(rng.uniform(*alpha_range))
        pops = np.sort(rng.choice(Kw, Kc, replace=False))
        Qt = np.asfortranarray(rng.dirichlet(alpha * np.ones(Kc), size=I).T)
        Pt = np.asfortranarray(P_world[pops, :J])
        ctx = HIPAdmixData(I, J, Kc, 3)
        try:
            ctx.generate_bed(Qt, Pt, seed * 1000 + c, miss_rate)
            d = fit_on_device(ctx, Kc, seed * 1000 + c + 7, max_iter=max_iter)
            F = panel.features(ctx)
        finally:
            ctx.close()
        perm, corr = match_components(d.p, Pt)
        Tc = np.zeros((I, Kw), np.float32)
        Tc[:, pops[perm]] = d.q.T
        X.append(F)
        T.append(Tc)
        info.append(dict(K=Kc, alpha=alpha, pops=pops.tolist(), iterations=d.iterations,
                         min_match_corr=float(corr.min())))
        if log:
            log(f"cohort {c}: K={Kc} alpha={alpha:.2f} {d.iterations} iterations, match corr >= {corr.min():.3f}")
    return np.concatenate(X), np.concatenate(T), info


def warm_started_fit(ctx: HIPAdmixData, net: MLP, panel: Panel, seed: int, max_iter: int = 1000,
                     rtol: float = 1e-7, p_steps: int = 2, judge: bool = True, em_iters: int = 5, Q: int = 3,
                     stop_at=None, min_gain: float = 1.0, q_steps: int = 1):
    """admixture_qn! from the network's proposal (with the EM judge/fallback).

    p starts at the observed allele frequencies and gets p_steps fixed-Q MM
    steps: update_p! with q_next = q (as admixture_qn! sets it under fix_q,
    src/algorithms_outer.jl:98-101), p = p_next after each, and one
    log-likelihood pass at the end.  Two steps reach the P of a fixed-Q
    quasi-Newton iteration (2 update_p! + 2 log-likelihoods + extrapolation)
    for one log-likelihood pass less; one step left the fit an iteration
    longer at I = J = 100,000, K = 8, and none doubled the iterations at
    I = 200,000, J = 50,000, K = 10.

    An accepted proposal then gets q_steps fixed-P steps: update_q! at (q, p),
    q = q_next, and one log-likelihood pass, the start of admixture_qn!.  One
    step took the fit at I = J = 100,000, K = 8 to |LL - LL*| < 1e-4 in one
    quasi-Newton iteration instead of two (the start's gap to LL* 1.3e-2 ->
    3.9e-3 at I = 200,000, J = 50,000, K = 10, where it still needs three).

    The judge keeps the proposal when its log-likelihood (after the P
    refinement) is at least that of one FRAPPE EM step from the random start
    and it explains the genotypes beyond Hardy-Weinberg: with p at the observed
    allele frequencies every q gives the same, structure-free log-likelihood
    LL_hw (s = q.f = f), and refining p for a q that carries no information
    gains about (K-1)J/2 by chance (half the chi-square mean of the (K-1)J free
    allele frequencies), so the gain LL_warm - LL_hw must exceed
    min_gain * (K-1) * J.  Otherwise the fit falls back to the random start.

    Returns dict(history=[(seconds, LL)], chosen='warm'|'em', ll_warm, ll_em,
    ll_hw, gain (LL_warm - LL_hw in units of (K-1)J), iterations).  history[0] is the starting point; seconds count from after
    the host AdmixData allocation (features, inference, P refinement and the
    judge included), the same origin as the plain run's timing.  With
    stop_at = (LL*, tol) the run ends at the first iteration within tol of
    LL*."""
    if net.out_dim != ctx.K:
        raise ValueError("network output must equal the context's K")
    I, J, K = ctx.I, ctx.J, ctx.K
    # host state (random QN history and the judge's random start) is allocated
    # before the clock starts, as for the plain run's AdmixData
    d = AdmixData(I, J, K, Q, rng=np.random.default_rng(seed), T=ctx.T)
    ctx.synchronize()
    t0 = time.perf_counter()
    panel.features(ctx, to_host=False)
    ctx.warm_q(net)
    # observed allele frequencies, on the device, and the log-likelihood there
    # (p_k = f for every k: s = f whatever q is), from the same counts
    ll_hw = ctx.warm_p()
    # admixture_qn! iterations from iter = 1 write every quasi-Newton history
    # column before reading it (src/algorithms_outer.jl:125-133), so the
    # device's U and V need no upload for these fresh runs
    ll_warm = ll_hw
    if p_steps > 0:
        ctx.copy_block(N.ADMX_X_NEXT, N.ADMX_X, N.ADMX_SIDE_Q)  # q_next = q: Q fixed
        for _ in range(p_steps):
            ctx.update_p()  # p_next from (q_next, p)
            ctx.copy_block(N.ADMX_X, N.ADMX_X_NEXT, N.ADMX_SIDE_P)
        ll_warm = ctx.loglikelihood(N.ADMX_X)
    ctx.set_ll(float("nan"), ll_warm)
    gain = (ll_warm - ll_hw) / max((K - 1) * J, 1)
    chosen, ll_em = "warm", None
    if judge:
        # the warm start waits in x_tmp while one FRAPPE EM step runs from the
        # random start (init_em!, src/algorithms_outer.jl:12-45)
        ctx.copy_x(N.ADMX_X_TMP, N.ADMX_X)
        ctx.set_x(N.ADMX_X, d.x)
        ctx.init_em(1)
        ll_em = ctx.get_ll()[1]
        if not (ll_warm >= ll_em and gain > min_gain):
            chosen = "em"
            if em_iters > 1:
                ctx.init_em(em_iters - 1)
        else:
            ctx.copy_x(N.ADMX_X, N.ADMX_X_TMP)
    ll0 = ll_warm if chosen == "warm" else ctx.get_ll()[1]
    if chosen == "warm" and q_steps > 0:
        for _ in range(q_steps):  # P fixed: q from update_q! at (q, p)
            ctx.update_q()
            ctx.copy_block(N.ADMX_X, N.ADMX_X_NEXT, N.ADMX_SIDE_Q)
        ll0 = ctx.loglikelihood(N.ADMX_X)
    ctx.set_ll(float("nan"), ll0)
    hist = [(time.perf_counter() - t0, ll0)]
    its = 0
    for it in range(1, max_iter + 1):
        r = ctx.qn_iteration(it)
        hist.append((time.perf_counter() - t0, r["new"]))
        its = it
        if r["reldiff"] < rtol:
            break
        if stop_at is not None and abs(r["new"] - stop_at[0]) <= stop_at[1] * abs(stop_at[0]):
            break
    return dict(history=hist, chosen=chosen, ll_warm=ll_warm, ll_em=ll_em, ll_hw=ll_hw, gain=gain, iterations=its,
                converged=its < max_iter and not math.isnan(hist[-1][1]))


__all__ = ["MLP", "Panel", "world_frequencies", "drifted_populations", "match_components", "training_set",
           "warm_started_fit"]
