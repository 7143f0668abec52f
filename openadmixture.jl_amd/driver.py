"""Driver entry (src/driver.jl:35-75, 129-206) on the HIP backend."""
from __future__ import annotations

import numpy as np

from .algorithms import admixture_qn, init_em
from .hip import HIPAdmixData
from .plink import SnpArray, SnpLinAlg
from .structs import AdmixData


def hip_admixture_base(d: AdmixData, g_la: SnpLinAlg, I: int, J: int, stream_snps: int = 0):
    """Counterpart of _cu_admixture_base (src/cuda/structs.jl:33-38): allocate
    the device mirror and copy the BED payload to HBM once (stream_snps > 0:
    keep it in pinned host memory and stream it through HBM in chunks of that
    many SNPs, for cohorts larger than HBM).  Returns (d_cu, g_cu); both are
    the same device context."""
    ctx = HIPAdmixData(I, J, d.K, d.Q, d.T, stream_snps=stream_snps)
    ctx.upload_bed(g_la.s.data)
    d.hip = ctx
    return ctx, ctx


def _admixture_base(filename, K, n_iter=1000, rtol=1e-7, rng=None, em_iters=5, T=np.float64, Q=3,
                    use_gpu=True, verbose=False, progress_bar=False, fix_q=False, fix_p=False,
                    init_q=None, init_p=None, log=print, mode="ZAL", stream_snps=0):
    if not use_gpu:
        raise ValueError("this package computes on the HIP device only (use_gpu=true)")
    if log:
        log("Loading genotype data...")
    g = SnpArray(filename)
    g_la = SnpLinAlg(g, T)
    I, J = g_la.shape
    if log:
        log(f"Loaded {I} samples and {J} SNPs")
    d = AdmixData(I, J, K, Q, skipmissing=True, rng=rng, T=T)
    d_cu, g_cu = hip_admixture_base(d, g_la, I, J, stream_snps=stream_snps)
    if init_q is None or init_p is None:
        init_em(d, g_la, em_iters, d_cu=d_cu, g_cu=g_cu, log=log)
    if init_q is not None:
        d.q[...] = init_q
    if init_p is not None:
        d.p[...] = init_p
    if init_q is not None or init_p is not None:
        # the reference leaves d.ll_new stale here (src/driver.jl:157-166); recompute it
        d_cu.set_x(0, d.x)
        d.ll_new = d_cu.loglikelihood(0)
    admixture_qn(d, g_la, n_iter, rtol, d_cu=d_cu, g_cu=g_cu, mode=mode, fix_q=fix_q, fix_p=fix_p,
                 verbose=verbose, log=log)
    return d


def run_admixture(filename, K, rng=None, sparsity=None, prefix=None, skfr_tries=1, skfr_max_inner_iter=50,
                  skfr_mode="global", admix_n_iter=1000, admix_rtol=1e-7, admix_n_em_iters=5, T=np.float64, Q=3,
                  use_gpu=True, verbose=False, progress_bar=False, fix_q=False, fix_p=False, init_q=None,
                  init_p=None, log=print, mode="ZAL", stream_snps=0):
    """run_admixture(filename, K; ...)  (src/driver.jl:35-75).  Returns
    (d, clusters, aims); the SKFR pre-selection (sparsity) is not provided."""
    if not filename.endswith(".bed"):
        raise ValueError("filename should end with .bed")
    if sparsity is not None:
        raise NotImplementedError("SKFR AIM pre-selection (sparsity=...) is outside this package")
    if log:
        log(f"Using {filename} as input.")
    d = _admixture_base(filename, K, n_iter=admix_n_iter, rtol=admix_rtol, rng=rng, em_iters=admix_n_em_iters,
                        T=T, Q=Q, use_gpu=use_gpu, verbose=verbose, progress_bar=progress_bar, fix_q=fix_q,
                        fix_p=fix_p, init_q=init_q, init_p=init_p, log=log, mode=mode, stream_snps=stream_snps)
    return d, None, None
