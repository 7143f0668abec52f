"""Launch-shape autotuner for the genotype passes.

The tunable is the number of waves a genotype pass is split into
(ADMX_TUNE_TARGET_WAVES): it fixes the SNP chunk per workgroup on the Q side
and the individual chunk on the P side, trading occupancy and tail balance
against the size of the per-chunk partial sums that the finalize kernels
reduce.  The tuner times one update_q! + update_p! pair per candidate with
HIP events inside the library and keeps the fastest; results are cached per
(I, J, K, precision) for the life of the process.
"""
from __future__ import annotations

from . import _native as N

CANDIDATES = (8192, 16384, 32768, 65536, 131072)
_cache: dict = {}


def autotune_waves(ctx, candidates=CANDIDATES, reps: int = 2) -> int:
    """Pick and set the fastest waves-per-pass for ctx (its state x must hold a
    valid point; the passes write x_next only).  Returns the chosen value."""
    prec = getattr(ctx, "_precision", "f64")
    key = (ctx.I, ctx.J, ctx.K, str(ctx.T), prec)
    if key in _cache:
        ctx.set_tuning(N.ADMX_TUNE_TARGET_WAVES, _cache[key])
        return _cache[key]
    best, best_ms = None, float("inf")
    ctx.profile(True)
    try:
        for w in candidates:
            ctx.set_tuning(N.ADMX_TUNE_TARGET_WAVES, w)
            ctx.update_q(False)
            ctx.update_p(False)
            ctx.profile_reset()
            for _ in range(reps):
                ctx.update_q(False)
                ctx.update_p(False)
            ms = sum(t for _, t in ctx.profile_report().values())
            if ms < best_ms:
                best, best_ms = w, ms
    finally:
        ctx.profile(False)
    ctx.set_tuning(N.ADMX_TUNE_TARGET_WAVES, best)
    _cache[key] = best
    return best
