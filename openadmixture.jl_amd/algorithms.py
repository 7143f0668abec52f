"""Inner and outer algorithms (src/algorithms_inner.jl, src/algorithms_outer.jl)
on the HIP backend.

The reference branches on ``d_cu === nothing`` between its CPU loops and its
CUDA runners; this package has a single backend, the HIP device context, so
``d_cu`` is required (or attached to ``d`` as ``d.hip``) and there is no CPU
path.  Host-level functions (em_q!, em_p!, update_q!, update_p!) move the
state between the host AdmixData and HBM around the device work, as the
reference's GPU branch copies q/p and XtX/Xtz; ``init_em`` and
``admixture_qn`` keep the whole iteration device-resident and only read back
the log-likelihoods each iteration and the state at the end.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import _native as N
from .hip import HIPAdmixData
from .loops import _ctx
from .structs import AdmixData


def _dev(d: AdmixData, d_cu) -> HIPAdmixData:
    c = d_cu if d_cu is not None else getattr(d, "hip", None)
    if not isinstance(c, HIPAdmixData):
        raise TypeError("a HIP device context is required (d_cu=..., from hip_admixture_base); "
                        "this package has no CPU path")
    if (c.I, c.J, c.K) != (d.I, d.J, d.K) or c.T != d.T:
        raise ValueError("device context does not match AdmixData (I, J, K, T)")
    # the reference's CUDA path asserts skipmissing (src/algorithms_inner.jl:191, 267);
    # this backend runs both loop families
    c.set_skipmissing(d.skipmissing)
    return c


def loglikelihood(g, q, p, qp_small=None, K=None, skipmissing=True) -> float:
    """loglikelihood(g, q, p, qp_small, K, skipmissing)  (src/algorithms_inner.jl:27-35):
    loglikelihood_loop_skipmissing or, with skipmissing=False, the mean-imputation
    loglikelihood_loop (src/loops.jl:493-514)."""
    c = _ctx(g)
    c.set_skipmissing(skipmissing)
    try:
        return c.loglikelihood_loop(np.asfortranarray(q), np.asfortranarray(p), 0, c.I, 0, c.J)
    finally:
        c.set_skipmissing(True)


def em_q(d: AdmixData, g=None, mode="base", d_cu=None, verbose=False) -> None:
    """em_q!(d, g; mode)  (src/algorithms_inner.jl:54-92): divided by 2J.
    mode "base" / "fallback": q_next from (q, p); "fallback2": q_next2 from
    (q_next, p_next)."""
    mode = str(mode).lstrip(":")
    c = _dev(d, d_cu)
    _sync_in(c, d)
    c.em_q(mode)
    if mode == "fallback2":
        d.x_next2[:, : d.I] = c.get_block(N.ADMX_X_NEXT2, N.ADMX_SIDE_Q)
    else:
        d.x_next[:, : d.I] = c.get_block(N.ADMX_X_NEXT, N.ADMX_SIDE_Q)


def em_p(d: AdmixData, g=None, mode="base", d_cu=None, verbose=False) -> None:
    """em_p!(d, g; mode)  (src/algorithms_inner.jl:94-141): p_next = p_tmp / (p_tmp + p_next).
    mode "base": p_next from (q, p); "fallback": p_next from (q_next, p);
    "fallback2": p_next2 from (q_next2, p_next)."""
    mode = str(mode).lstrip(":")
    c = _dev(d, d_cu)
    _sync_in(c, d)
    c.em_p(mode)
    if mode == "fallback2":
        d.x_next2[:, d.I:] = c.get_block(N.ADMX_X_NEXT2, N.ADMX_SIDE_P)
    else:
        d.x_next[:, d.I:] = c.get_block(N.ADMX_X_NEXT, N.ADMX_SIDE_P)


def _sync_in(c: HIPAdmixData, d: AdmixData):
    c.set_x(N.ADMX_X, d.x)
    c.set_x(N.ADMX_X_NEXT, d.x_next)
    c.set_x(N.ADMX_X_NEXT2, d.x_next2)


def update_q(d: AdmixData, g=None, update2=False, d_cu=None, g_cu=None, verbose=False):
    """update_q!(d, g, update2)  (src/algorithms_inner.jl:167-231): gradient and
    Hessian pass, Xtz .*= -1, batched simplex QPs, q_next = q + delta,
    project_q!.  Returns the QP return codes."""
    c = _dev(d, d_cu)
    d.ll_prev = d.ll_new
    _sync_in(c, d)
    status = c.update_q(update2, status=True)
    tgt = N.ADMX_X_NEXT2 if update2 else N.ADMX_X_NEXT
    out = c.get_x(tgt)
    dst = d.x_next2 if update2 else d.x_next
    dst[:, : d.I] = out[:, : d.I]
    d.XtX_q[...], d.Xtz_q[...] = c.hessians(N.ADMX_SIDE_Q)
    return status


def update_p(d: AdmixData, g=None, update2=False, d_cu=None, g_cu=None, verbose=False):
    """update_p!(d, g, update2)  (src/algorithms_inner.jl:245-310): uses q_next
    (Gauss-Seidel), batched box QPs, p_next = p + delta, project_p!."""
    c = _dev(d, d_cu)
    d.ll_prev = d.ll_new
    _sync_in(c, d)
    status = c.update_p(update2, status=True)
    tgt = N.ADMX_X_NEXT2 if update2 else N.ADMX_X_NEXT
    out = c.get_x(tgt)
    dst = d.x_next2 if update2 else d.x_next
    dst[:, d.I:] = out[:, d.I:]
    d.XtX_p[...], d.Xtz_p[...] = c.hessians(N.ADMX_SIDE_P)
    return status


def init_em(d: AdmixData, g=None, iter: int = 5, d_cu=None, g_cu=None, verbose=False, log=print):
    """init_em!(d, g, iter)  (src/algorithms_outer.jl:12-45): FRAPPE EM priming."""
    c = _dev(d, d_cu)
    if log:
        log(f"Performing {iter} EM steps for priming")
    c.set_x(N.ADMX_X, d.x)
    t0 = time.perf_counter()
    lls = c.init_em(iter)
    dt = (time.perf_counter() - t0) / max(iter, 1)
    if log:
        for i, ll in enumerate(lls, 1):
            log(f"EM Iteration {i} ({dt} sec): Loglikelihood = {ll}")
    c.get_x(N.ADMX_X, d.x)
    d.ll_new = c.get_ll()[1]
    return lls


def admixture_qn(d: AdmixData, g=None, iter: int = 1000, rtol: float = 1e-7, d_cu=None, g_cu=None,
                 mode="ZAL", iter_count_offset=0, fix_p=False, fix_q=False, verbose=False,
                 progress_bar=False, log=print):
    """admixture_qn!(d, g, iter, rtol; mode=:ZAL)  (src/algorithms_outer.jl:63-193).

    Each iteration: two SQP map applications (update_q!, update_p! twice),
    LL of the basic point, ZAL (mode="ZAL") or limited-memory Broyden
    (mode="LBQN") extrapolation with Q history columns,
    projections, LL of the QN point, accept the QN point iff ll_prev < ll_qn,
    stop when |ll_new - ll_prev| / |ll_prev| < rtol.  Returns the LL history.
    """
    mode = str(mode).lstrip(":")
    if mode not in ("ZAL", "LBQN"):
        raise ValueError(f"Invalid mode {mode!r}")
    c = _dev(d, d_cu)
    c.set_qn_mode(mode)
    c.set_x(N.ADMX_X, d.x)
    c.set_uv(d.U, d.V)
    if math.isnan(d.ll_new):
        d.ll_new = c.loglikelihood(N.ADMX_X)
    c.set_ll(d.ll_prev, d.ll_new)
    if log:
        log(f"initial ll: {d.ll_new}")
        log("Starting main algorithm")
    llhist = [d.ll_new]
    converged = False
    i = iter_count_offset
    t_all = time.perf_counter()
    for i in range(iter_count_offset + 1, iter + 1):
        t0 = time.perf_counter()
        r = c.qn_iteration(i, fix_q, fix_p)
        dt = time.perf_counter() - t0
        if log:
            log(f"Iteration {i} ({dt} sec): LogLikelihood = {r['new']}, reldiff = {r['reldiff']}")
            log(f"    LogLikelihoods: Previous = {r['prev']} QN = {r['qn']}, Basic = {r['basic']}")
        llhist.append(r["new"])
        if r["reldiff"] < rtol:
            converged = True
            break
    total = time.perf_counter() - t_all
    if log:
        if converged:
            log(f"Main algorithm converged in {i} iterations over {total} sec.")
        else:
            log(f"Main algorithm failed to converge after {i} iterations over {total} sec.")
    for which, arr in ((N.ADMX_X, d.x), (N.ADMX_X_NEXT, d.x_next), (N.ADMX_X_NEXT2, d.x_next2),
                       (N.ADMX_X_TMP, d.x_tmp)):
        c.get_x(which, arr)
    c.get_uv(d.U, d.V)
    d.ll_prev, d.ll_new = c.get_ll()
    d.llhist = llhist
    d.converged = converged
    d.iterations = i
    return llhist
