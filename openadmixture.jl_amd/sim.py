"""Synthetic admixed cohorts (the workloads the path is measured on).

True parameters are drawn on the host: Q[:, i] ~ Dirichlet(alpha 1_K) and
P[k, j] ~ Uniform(lo, hi) or Beta(a, a) clipped to [lo, hi].  Genotypes
g_ij ~ Binomial(2, q_i . p_j), each entry missing with probability
``miss_rate``, are generated on the device straight into the packed 2-bit BED
(a seeded counter hash, so the same (Q, P, seed) give the same bytes on any
device or in the CPU oracle's generator).
"""
from __future__ import annotations

import numpy as np

from .hip import HIPAdmixData


def draw_truth(I: int, J: int, K: int, alpha: float = 0.5, p_dist: str = "uniform", p_lo: float = 0.05,
               p_hi: float = 0.95, beta_a: float = 0.5, seed: int = 1):
    rng = np.random.default_rng(seed)
    Qt = np.asfortranarray(rng.dirichlet(alpha * np.ones(K), size=I).T)
    if p_dist == "uniform":
        Pt = rng.uniform(p_lo, p_hi, size=(K, J))
    elif p_dist == "beta":
        Pt = np.clip(rng.beta(beta_a, beta_a, size=(K, J)), p_lo, p_hi)
    else:
        raise ValueError(p_dist)
    return Qt, np.asfortranarray(Pt)


# The five workloads of the benchmark plan: (I, J, K, alpha, P law, missing rate).
CONFIGS = {
    "c1": dict(I=1_000, J=10_000, K=3, alpha=0.5, p_dist="uniform", p_lo=0.05, p_hi=0.95, miss_rate=0.01),
    "c2": dict(I=10_000, J=100_000, K=5, alpha=0.3, p_dist="uniform", p_lo=0.05, p_hi=0.95, miss_rate=0.02),
    "c3": dict(I=100_000, J=100_000, K=8, alpha=0.3, p_dist="beta", p_lo=0.01, p_hi=0.99, miss_rate=0.02),
    "c4": dict(I=20_000, J=200_000, K=6, alpha=0.3, p_dist="beta", p_lo=0.01, p_hi=0.99, miss_rate=0.02),
}


def cohort_on_device(ctx: HIPAdmixData, alpha=0.5, p_dist="uniform", p_lo=0.05, p_hi=0.95, miss_rate=0.01,
                     seed=1, K_true=None):
    """Draw truth (K_true populations, default ctx.K) and generate the BED in HBM."""
    K = ctx.K if K_true is None else K_true
    if K != ctx.K:
        raise ValueError("the device generator needs K_true == ctx.K")
    Qt, Pt = draw_truth(ctx.I, ctx.J, K, alpha, p_dist, p_lo, p_hi, seed=seed)
    ctx.generate_bed(Qt, Pt, seed, miss_rate)
    return Qt, Pt
