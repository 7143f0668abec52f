"""The tile loops of src/loops.jl with their reference signatures, executed by
the HIP kernels.

Each function keeps the argument order of its Julia counterpart
(``f!(outs..., g, q, p, [p_tmp,] qp_small, irange, jrange, K)``) and its
accumulate (``+=``) semantics into the caller's arrays; callers zero the
outputs and apply the post-scaling, exactly as algorithms_inner.jl does.
``g`` is the device genotype handle returned by ``hip_admixture_base`` (the
counterpart of the reference's ``g_cu``); ``qp_small`` is accepted and ignored
(the kernels keep s = q_i . p_j in registers).  Ranges are ``UnitRange(a, b)``
(1-based, inclusive, as Julia's a:b) or Python ``range`` objects (0-based).
"""
from __future__ import annotations

from .hip import HIPAdmixData


class UnitRange:
    """Julia's a:b (1-based, inclusive)."""

    def __init__(self, first: int, last: int):
        self.first, self.last = int(first), int(last)

    def __len__(self):
        return max(0, self.last - self.first + 1)

    def __repr__(self):
        return f"{self.first}:{self.last}"


def _bounds(r):
    if isinstance(r, UnitRange):
        return r.first - 1, r.last
    if isinstance(r, range):
        if r.step != 1:
            raise ValueError("ranges must be contiguous")
        return r.start, r.stop
    a, b = r
    return int(a), int(b)


def _ctx(g) -> HIPAdmixData:
    if isinstance(g, HIPAdmixData):
        return g
    h = getattr(g, "hip", None)
    if isinstance(h, HIPAdmixData):
        return h
    raise TypeError("g must be the device genotype handle from hip_admixture_base(d, g, I, J)")


def _check_k(ctx, K):
    if K is not None and int(K) != ctx.K:
        raise ValueError(f"K={K} does not match the device context (K={ctx.K})")


def loglikelihood_loop_skipmissing(g, q, p, qp_small, irange, jrange, K) -> float:
    """src/loops.jl:517-539: sum over the tile of m*(g log s + (2-g) log(1-s)), Float64."""
    c = _ctx(g)
    _check_k(c, K)
    i0, i1 = _bounds(irange)
    j0, j1 = _bounds(jrange)
    return c.loglikelihood_loop(q, p, i0, i1, j0, j1)


def em_q_loop_skipmissing(q_next, g, q, p, qp_small, irange, jrange, K) -> None:
    """src/loops.jl:610-632: q_next[k,i] += m*(g q p/s + (2-g) q (1-p)/(1-s))."""
    c = _ctx(g)
    _check_k(c, K)
    i0, i1 = _bounds(irange)
    j0, j1 = _bounds(jrange)
    c.em_q_loop(q_next, q, p, i0, i1, j0, j1)


def em_p_loop_skipmissing(p_next, g, q, p, p_tmp, qp_small, irange, jrange, K) -> None:
    """src/loops.jl:700-721: p_tmp += m g q p/s ; p_next += m (2-g) q (1-p)/(1-s)."""
    c = _ctx(g)
    _check_k(c, K)
    i0, i1 = _bounds(irange)
    j0, j1 = _bounds(jrange)
    c.em_p_loop(p_next, p_tmp, q, p, i0, i1, j0, j1)


def update_q_loop_skipmissing(XtX, Xtz, g, q, p, qp_small, irange, jrange, K) -> None:
    """src/loops.jl:867-894: gradient and K x K Hessian of the Q side (ADMIXTURE form)."""
    c = _ctx(g)
    _check_k(c, K)
    i0, i1 = _bounds(irange)
    j0, j1 = _bounds(jrange)
    c.update_q_loop(XtX, Xtz, q, p, i0, i1, j0, j1)


def update_p_loop_skipmissing(XtX, Xtz, g, q, p, qp_small, irange, jrange, K) -> None:
    """src/loops.jl:923-950: gradient and K x K Hessian of the P side."""
    c = _ctx(g)
    _check_k(c, K)
    i0, i1 = _bounds(irange)
    j0, j1 = _bounds(jrange)
    c.update_p_loop(XtX, Xtz, q, p, i0, i1, j0, j1)


# ---------------------------------------------------------------- mean-imputation loops
def _impute(c, fn, *args):
    c.set_skipmissing(False)
    try:
        return fn(*args)
    finally:
        c.set_skipmissing(True)


def loglikelihood_loop(g, q, p, qp_small, irange, jrange, K) -> float:
    """src/loops.jl:493-514: sum of g log s + (2-g) log(1-s) with a missing g = mu_j."""
    c = _ctx(g)
    _check_k(c, K)
    return _impute(c, c.loglikelihood_loop, q, p, *_bounds(irange), *_bounds(jrange))


def em_q_loop(q_next, g, q, p, qp_small, irange, jrange, K) -> None:
    """src/loops.jl:589-608: em_q_loop! (missing g = mu_j)."""
    c = _ctx(g)
    _check_k(c, K)
    _impute(c, c.em_q_loop, q_next, q, p, *_bounds(irange), *_bounds(jrange))


def em_p_loop(p_next, g, q, p, p_tmp, qp_small, irange, jrange, K) -> None:
    """src/loops.jl:679-698: em_p_loop! (missing g = mu_j)."""
    c = _ctx(g)
    _check_k(c, K)
    _impute(c, c.em_p_loop, p_next, p_tmp, q, p, *_bounds(irange), *_bounds(jrange))


def update_q_loop(XtX, Xtz, g, q, p, qp_small, irange, jrange, K) -> None:
    """src/loops.jl:840-865: update_q_loop! (missing g = mu_j)."""
    c = _ctx(g)
    _check_k(c, K)
    _impute(c, c.update_q_loop, XtX, Xtz, q, p, *_bounds(irange), *_bounds(jrange))


def update_p_loop(XtX, Xtz, g, q, p, qp_small, irange, jrange, K) -> None:
    """src/loops.jl:896-921: update_p_loop! (missing g = mu_j)."""
    c = _ctx(g)
    _check_k(c, K)
    _impute(c, c.update_p_loop, XtX, Xtz, q, p, *_bounds(irange), *_bounds(jrange))
