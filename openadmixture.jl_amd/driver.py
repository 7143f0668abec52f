"""Driver entry (src/driver.jl:35-75, 129-206) on the HIP backend."""
from __future__ import annotations

import numpy as np

from .algorithms import admixture_qn, init_em
from .hip import HIPAdmixData
from .plink import SnpArray, SnpLinAlg, filter_plink
from .skfr import ImputedSnpMatrix, sparsekmeans1, sparsekmeans2, sparsekmeans_path, sparsekmeans_repeat
from .structs import AdmixData


def hip_admixture_base(d: AdmixData, g_la: SnpLinAlg, I: int, J: int, stream_snps: int = 0,
                       stream_from_file: bool = False, resident_snps: int = 0):
    """Counterpart of _cu_admixture_base (src/cuda/structs.jl:33-38): allocate
    the device mirror and copy the BED payload to HBM once.  stream_snps > 0
    streams it through HBM in chunks of that many SNPs instead, for cohorts
    larger than HBM: from pinned host memory, or with stream_from_file from
    the .bed file itself (cohorts larger than host memory); resident_snps
    leading SNPs then stay in HBM.  Returns (d_cu, g_cu); both are the same
    device context."""
    if stream_from_file:
        if not stream_snps:
            raise ValueError("stream_from_file needs stream_snps > 0")
        ctx = HIPAdmixData(I, J, d.K, d.Q, d.T, stream_snps=stream_snps, bed_file=g_la.s.path,
                           resident_snps=resident_snps)
    else:
        ctx = HIPAdmixData(I, J, d.K, d.Q, d.T, stream_snps=stream_snps)
        ctx.upload_bed(g_la.s.data)
        if resident_snps:
            ctx.set_resident_snps(resident_snps)
    d.hip = ctx
    return ctx, ctx


def _admixture_base(filename, K, n_iter=1000, rtol=1e-7, rng=None, em_iters=5, T=np.float64, Q=3,
                    use_gpu=True, verbose=False, progress_bar=False, fix_q=False, fix_p=False,
                    init_q=None, init_p=None, log=print, mode="ZAL", stream_snps=0, stream_from_file=False,
                    resident_snps=0):
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
    d_cu, g_cu = hip_admixture_base(d, g_la, I, J, stream_snps=stream_snps, stream_from_file=stream_from_file,
                                    resident_snps=resident_snps)
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


def _filter_SKFR(filename, K, sparsity, rng=None, prefix=None, tries=10, max_inner_iter=50, ftn=sparsekmeans1,
                 log=print, nsplit=0):
    """_filter_SKFR(filename, K, sparsity; rng, prefix, tries, max_inner_iter, ftn)
    (src/driver.jl:80-101; the vector-of-sparsities method :106-123): SKFR on
    the device, then the PLINK trio reduced to the AIMs (sorted SNP order) as
    prefix_K_{sparsity}aims.  Returns (bed path, clusters, aims) for one
    sparsity, (bed paths, clusters per level, aims per level) for a list."""
    if not filename.endswith(".bed"):
        raise ValueError("filename should end with .bed")
    prefix = filename[:-4] if prefix is None else prefix
    X = ImputedSnpMatrix(filename, K, rng=rng, nsplit=nsplit)
    if isinstance(sparsity, (list, tuple, np.ndarray)):
        clusters, aims = sparsekmeans_path(X, sparsity, iter=tries, max_inner_iter=max_inner_iter, ftn=ftn)
        outs = []
        for s, a in zip(sparsity, aims):
            des = f"{prefix}_{K}_{int(s)}aims"
            filter_plink(filename, None, np.sort(a), des)
            outs.append(des)
        return outs, clusters, aims
    if tries == 1:
        clusters, _, aims, _, _ = ftn(X, int(sparsity), max_iter=max_inner_iter, squares=False)
    else:
        clusters, _, aims, _, _, _ = sparsekmeans_repeat(X, int(sparsity), iter=tries, max_inner_iter=max_inner_iter,
                                                         ftn=ftn)
    des = f"{prefix}_{K}_{int(sparsity)}aims"
    if log:
        log(des)
    filter_plink(filename, None, np.sort(aims), des)
    return des + ".bed", clusters, aims


def run_admixture(filename, K, rng=None, sparsity=None, prefix=None, skfr_tries=1, skfr_max_inner_iter=50,
                  skfr_mode="global", admix_n_iter=1000, admix_rtol=1e-7, admix_n_em_iters=5, T=np.float64, Q=3,
                  use_gpu=True, verbose=False, progress_bar=False, fix_q=False, fix_p=False, init_q=None,
                  init_p=None, log=print, mode="ZAL", stream_snps=0, stream_from_file=False, resident_snps=0):
    """run_admixture(filename, K; ...)  (src/driver.jl:35-75).  Returns
    (d, clusters, aims); with sparsity given, SKFR first selects that many
    AIMs (skfr_mode "global": sparsekmeans1, "local": sparsekmeans2) and the
    fit runs on the filtered PLINK file prefix_K_{sparsity}aims.bed.  Cluster
    labels and AIM indices are 0-based.  stream_snps / stream_from_file /
    resident_snps select J-chunk streaming (see hip_admixture_base)."""
    if not filename.endswith(".bed"):
        raise ValueError("filename should end with .bed")
    if log:
        log(f"Using {filename} as input.")
    if sparsity is not None:
        mode_s = str(skfr_mode).lstrip(":")
        if mode_s == "global":
            ftn = sparsekmeans1
        elif mode_s == "local":
            ftn = sparsekmeans2
        else:
            raise ValueError("skfr_mode can only be 'global' or 'local'")
        admix_input, clusters, aims = _filter_SKFR(filename, K, sparsity, rng=rng, prefix=prefix,
                                                   tries=skfr_tries, max_inner_iter=skfr_max_inner_iter, ftn=ftn,
                                                   log=log)
    else:
        admix_input, clusters, aims = filename, None, None
    d = _admixture_base(admix_input, K, n_iter=admix_n_iter, rtol=admix_rtol, rng=rng, em_iters=admix_n_em_iters,
                        T=T, Q=Q, use_gpu=use_gpu, verbose=verbose, progress_bar=progress_bar, fix_q=fix_q,
                        fix_p=fix_p, init_q=init_q, init_p=init_p, log=log, mode=mode, stream_snps=stream_snps,
                        stream_from_file=stream_from_file, resident_snps=resident_snps)
    return d, clusters, aims
