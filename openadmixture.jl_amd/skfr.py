"""Sparse k-means with feature ranking (SKFR) on the HIP device: the AIM
pre-selection of ``run_admixture(...; sparsity)`` (src/driver.jl:56-66,
80-123).

The reference calls SparseKmeansFeatureRanking.jl 0.1 for this step; the
names here mirror the calls it makes (``ImputedSnpMatrix{Float64}(g, K; rng)``,
``sparsekmeans1``, ``sparsekmeans2``, ``sparsekmeans_repeat``,
``sparsekmeans_path``).  Everything runs in libadmixhip's SKFR kernels on the
device-resident BED: per iteration one pass over all SNPs for the cluster
counts, the SNP ranking (device radix sort), and one pass over the selected
SNPs for the nearest-center reassignment.  Labels and SNP indices are
0-based.  Returns mirror the reference's tuples: ``(clusters, centers, aims,
WSS, obj)``; ``aims`` lists the selected SNPs in decreasing importance.
"""
from __future__ import annotations

import numpy as np

from .hip import HIPAdmixData
from .plink import SnpArray


def _seed(rng) -> int:
    if rng is None:
        return int(np.random.default_rng().integers(0, 2**63 - 1))
    if isinstance(rng, (int, np.integer)):
        return int(rng)
    return int(rng.integers(0, 2**63 - 1))


class ImputedSnpMatrix:
    """Genotypes on the device (standardized and mean-imputed inside the SKFR
    kernels) with k clusters initialised by k-means++ (the constructor of
    ``ImputedSnpMatrix{Float64}(g, K; rng)``, src/driver.jl:88).

    ``g`` is a ``.bed`` path, a :class:`SnpArray` or a :class:`HIPAdmixData`
    with the genotypes uploaded.  ``nsplit`` fixes the number of SNP slices of
    the distance sums (0: automatic); results are reproducible for a given
    value.
    """

    def __init__(self, g, k: int, rng=None, nsplit: int = 0):
        if isinstance(g, str):
            g = SnpArray(g)
        if isinstance(g, SnpArray):
            ctx = HIPAdmixData(g.I, g.J, int(k), 1, np.float64)
            ctx.upload_bed(g.data)
        elif isinstance(g, HIPAdmixData):
            ctx = g
        else:
            raise TypeError("g must be a .bed path, a SnpArray or a HIPAdmixData")
        self.ctx = ctx
        self.k = int(k)
        self.nsplit = int(nsplit)
        self.rng = rng
        self.clusters = ctx.skfr_init(self.k, _seed(rng), self.nsplit)

    @property
    def shape(self):
        return (self.ctx.I, self.ctx.J)

    def reinitialize(self, seed=None):
        """New k-means++ start (the restarts of sparsekmeans_repeat)."""
        self.clusters = self.ctx.skfr_init(self.k, _seed(self.rng if seed is None else seed), self.nsplit)
        return self.clusters


def _run(X: ImputedSnpMatrix, sparsity: int, mode: str, max_iter: int):
    cl, cen, aims, wss, obj, iters = X.ctx.skfr_run(X.k, X.clusters, int(sparsity), mode=mode,
                                                    max_iter=int(max_iter), nsplit=X.nsplit)
    X.clusters = cl
    X.iterations = iters
    return cl, cen, aims, wss, obj


def sparsekmeans1(X: ImputedSnpMatrix, sparsity: int, max_iter: int = 1000, squares: bool = False):
    """Global feature ranking: one list of `sparsity` SNPs for all clusters.
    Starts from X.clusters and leaves the result there.  ``squares`` is
    accepted for the reference's call signature (src/driver.jl:90) and has no
    effect.  Returns (clusters, centers (k x J), aims, WSS (k), obj)."""
    return _run(X, sparsity, "global", max_iter)


def sparsekmeans2(X: ImputedSnpMatrix, sparsity: int, max_iter: int = 1000, squares: bool = False):
    """Local feature ranking: `sparsity` SNPs per cluster; aims is their union
    ranked by the largest per-cluster score."""
    return _run(X, sparsity, "local", max_iter)


def sparsekmeans_repeat(X: ImputedSnpMatrix, sparsity: int, iter: int = 20, max_inner_iter: int = 20,
                        ftn=sparsekmeans1):
    """`iter` runs of ``ftn`` from independent k-means++ starts (the first
    from X's current clusters); keeps the run with the smallest objective
    (src/driver.jl:92).  Returns (clusters, centers, aims, WSS, obj, objs)
    with objs the objective of every run."""
    best, objs = None, []
    start = X.clusters.copy()
    for t in range(int(iter)):
        if t == 0:
            X.clusters = start.copy()
        else:
            X.reinitialize()
        res = ftn(X, sparsity, max_iter=max_inner_iter)
        objs.append(res[4])
        if best is None or res[4] < best[4]:
            best = res
    X.clusters = best[0].copy()
    return (*best, np.asarray(objs))


def sparsekmeans_path(X: ImputedSnpMatrix, sparsities, iter: int = 20, max_inner_iter: int = 20,
                      ftn=sparsekmeans1):
    """SKFR along decreasing sparsities (src/driver.jl:113-116): each level
    runs sparsekmeans_repeat with its first run started from the previous
    level's best clusters.  Returns (clusters per level, aims per level)."""
    sparsities = [int(s) for s in sparsities]
    if sparsities != sorted(sparsities, reverse=True):
        raise ValueError("sparsities should be decreasing")
    clusters, aims = [], []
    for s in sparsities:
        res = sparsekmeans_repeat(X, s, iter=iter, max_inner_iter=max_inner_iter, ftn=ftn)
        clusters.append(res[0])
        aims.append(res[2])
    return clusters, aims
