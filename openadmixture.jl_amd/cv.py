"""Uneven K sweep with n-fold cross-validation on one device-resident cohort.

The reference fits one K per ``run_admixture`` call (src/driver.jl:35-75) and
has no cross-validation; this module is the batched sweep built on the same
device primitives.  For every K it fits the full data and each of the
``nfolds`` holdout problems: fold f is the cohort with the entries of fold f
(a seeded hash partition of all genotype entries, ~1/nfolds each) set to
missing, fitted by EM priming + admixture_qn! like the full problem.  The CV
error of K is the held-out binomial deviance per observed held-out entry,

    CV(K) = sum_f sum_{(i,j) in fold f, observed} 2[g log(g/2s) + (2-g) log((2-g)/2(1-s))]
            / (number of such entries),   s = q_i . p_j of fold f's fit,

which is ADMIXTURE's deviance-residual CV criterion.  One context per K holds
that K's state; the fold genotypes are re-derived on the device from the
source context's BED (``admx_bed_from``), so the cohort is read from the host
once.  By default each fold fit starts from the full fit's estimates
(``fold_start="full"``): a fold hides only 1/nfolds of the entries, so its
optimum lies close to the full one and the fit needs a few iterations instead
of a random start's EM priming and full descent; ``fold_start="random"`` gives
every fold its own seeded random start and EM priming.
"""
from __future__ import annotations

import time

import numpy as np

from . import _native as N
from .algorithms import admixture_qn, init_em
from .hip import HIPAdmixData
from .structs import AdmixData
from .tune import autotune as _autotune


def fit_on_device(ctx: HIPAdmixData, K: int, seed: int, em_iters: int = 5, max_iter: int = 1000,
                  rtol: float = 1e-7, Q: int = 3):
    """EM priming + admixture_qn! on ctx's genotypes from a seeded random start;
    returns the host AdmixData (q, p, LL history, iteration count)."""
    d = AdmixData(ctx.I, ctx.J, K, Q, rng=np.random.default_rng(seed), T=ctx.T)
    init_em(d, None, em_iters, d_cu=ctx, log=None)
    admixture_qn(d, None, max_iter, rtol, d_cu=ctx, log=None)
    return d


def fit_from(ctx: HIPAdmixData, x, seed: int, max_iter: int = 1000, rtol: float = 1e-7, Q: int = 3):
    """admixture_qn! on ctx's genotypes from the state x (K x (I+J)), without
    EM priming; returns the host AdmixData."""
    d = AdmixData(ctx.I, ctx.J, ctx.K, Q, rng=np.random.default_rng(seed), T=ctx.T)
    d.x[...] = x
    admixture_qn(d, None, max_iter, rtol, d_cu=ctx, log=None)
    return d


def cv_sweep(src: HIPAdmixData, Ks, nfolds: int = 5, fold_seed: int = 0, init_seed: int = 1, em_iters: int = 5,
             max_iter: int = 1000, rtol: float = 1e-7, precision: str = "f64", Q: int = 3, autotune: bool = True,
             fold_start: str = "full", log=None, workers: int = 1, graphs: bool = True):
    """Fit K in Ks on src's genotypes, each with nfolds holdout problems.

    graphs: steady-state quasi-Newton iterations replay captured hipGraphs
    (ADMX_TUNE_GRAPHS; bit-identical to eager iterations).

    workers > 1 runs that many K values at once, each in its own context (own
    HIP stream) driven by its own host thread; the C ABI releases the
    interpreter lock for every call, so the problems' kernels overlap on the
    device.  Each K's results are those of the serial sweep (same contexts,
    same seeds, deterministic passes).

    Returns one dict per K (in the order of Ks): K, seconds (all nfolds + 1 fits
    of that K), the full fit's final LL and iteration count, the CV error and
    per-fold deviances, and the autotuner's choice (waves per pass, Hessian
    split per side)."""
    if nfolds < 2:
        raise ValueError("nfolds must be >= 2")
    if fold_start not in ("full", "random"):
        raise ValueError("fold_start must be 'full' or 'random'")
    Ks = [int(k) for k in Ks]
    args = (src, nfolds, fold_seed, init_seed, em_iters, max_iter, rtol, precision, Q, autotune, fold_start, log,
            graphs)
    if workers > 1 and len(Ks) > 1:
        from concurrent.futures import ThreadPoolExecutor

        if autotune:
            # tune each K alone first: timings taken while other problems run
            # would be noisy, and the cached choice keeps every K's chunking,
            # hence its results, those of the serial sweep
            for K in Ks:
                with HIPAdmixData(src.I, src.J, K, Q, src.T) as ctx:
                    ctx.set_precision(precision)
                    ctx.bed_from(src, -1)
                    ctx.set_x(N.ADMX_X, AdmixData(src.I, src.J, K, Q, rng=np.random.default_rng(init_seed), T=src.T).x)
                    _autotune(ctx)
        with ThreadPoolExecutor(max_workers=int(workers)) as ex:
            return list(ex.map(lambda K: _sweep_one(K, *args), Ks))
    return [_sweep_one(K, *args) for K in Ks]


def _sweep_one(K, src, nfolds, fold_seed, init_seed, em_iters, max_iter, rtol, precision, Q, tune, fold_start,
               log, graphs=True):
    """The full fit and the nfolds holdout fits of one K (one context)."""
    t0 = time.perf_counter()
    ctx = HIPAdmixData(src.I, src.J, int(K), Q, src.T)
    try:
        ctx.set_precision(precision)
        ctx.set_graphs(graphs)
        ctx.bed_from(src, -1)
        tuning = None
        if tune:
            ctx.set_x(N.ADMX_X, AdmixData(src.I, src.J, K, Q, rng=np.random.default_rng(init_seed), T=src.T).x)
            tuning = _autotune(ctx)
        full = fit_on_device(ctx, K, init_seed + 1000 * K, em_iters, max_iter, rtol, Q)
        devs, cnts, iters = [], [], []
        for f in range(nfolds):
            ctx.bed_from(src, f, nfolds, fold_seed)
            if fold_start == "full":
                d = fit_from(ctx, full.x, init_seed + 1000 * K + 1 + f, max_iter, rtol, Q)
            else:
                d = fit_on_device(ctx, K, init_seed + 1000 * K + 1 + f, em_iters, max_iter, rtol, Q)
            dev, cnt = ctx.heldout_deviance(src, f, nfolds, fold_seed)
            devs.append(dev)
            cnts.append(cnt)
            iters.append(d.iterations)
        ctx.synchronize()
    finally:
        ctx.close()
    el = time.perf_counter() - t0
    rec = dict(K=int(K), seconds=el, ll_full=float(full.ll_new), iterations_full=int(full.iterations),
               converged_full=bool(full.converged), iterations_folds=iters,
               cv_error=float(np.sum(devs) / max(np.sum(cnts), 1)),
               fold_deviance=[float(v) for v in devs], fold_entries=[int(v) for v in cnts],
               waves=tuning["waves"] if tuning else None, tuning=tuning)
    if log:
        log(f"K={K}: CV error {rec['cv_error']:.6f}, full LL {rec['ll_full']:.6e}, "
            f"{rec['iterations_full']} iterations, {el:.2f} s")
    return rec


def best_k(results):
    """K with the smallest CV error."""
    return min(results, key=lambda r: r["cv_error"])["K"]


__all__ = ["cv_sweep", "fit_on_device", "fit_from", "best_k"]
