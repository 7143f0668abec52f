"""Launch-shape autotuner for the genotype passes.

Two tunables, timed with HIP events inside the library:

* waves per pass (ADMX_TUNE_TARGET_WAVES): the SNP chunk per workgroup on the
  Q side and the individual chunk on the P side, trading occupancy and tail
  balance against the size of the per-chunk partial sums that the finalize
  kernels reduce;
* for K >= 9 (FP64 passes), whether the Hessian triangle runs in one launch
  or in two of half the accumulators each (ADMX_TUNE_SPLIT_Q / _SPLIT_P):
  two launches hold fewer registers per lane (more waves per SIMD) but read
  the genotypes and compute s = q.p twice;
* for the BF16 form, which kernel runs each side's update pass
  (ADMX_TUNE_QSIDE_MFMA / _PSIDE_MFMA): s, the Hessian and the gradient all on
  the matrix cores, or only the Hessian (s and the gradient as FP32 FMAs).
  The first does fewer vector instructions per genotype, the second fewer
  matrix instructions; which is faster depends on K (the default switches at
  K = 7, where they cross at I = J = 100,000).

The waves are tuned on update_q! + update_p! pairs with the default split
and kernel rules; then, at the chosen waves, each side's split (FP64,
K >= 9) or kernel (BF16) is tuned on that side's update alone.  Results are
cached per (I, J, K, dtype, precision) for the life of the process.  The
waves and the split give the same results up to the order of the per-chunk
partial sums (the split none at all); the two BF16 kernels agree within the
BF16 form's error bars against the FP64 oracle.
"""
from __future__ import annotations

from . import _native as N

CANDIDATES = (8192, 16384, 32768, 65536, 131072)
_cache: dict = {}


def _time(ctx, calls, reps):
    for f in calls:  # one untimed round: buffers grown, code loaded
        f()
    ctx.profile_reset()
    for _ in range(reps):
        for f in calls:
            f()
    return sum(t for _, t in ctx.profile_report().values())


def autotune(ctx, candidates=CANDIDATES, reps: int = 2) -> dict:
    """Pick and set the fastest waves per pass and, per side, the FP64 Hessian
    split (K >= 9) or the BF16 kernel for ctx (its state x must hold a valid
    point; the passes write x_next only).  Returns dict(waves, split_q,
    split_p, mfma_q, mfma_p) (-1: default rule)."""
    prec = getattr(ctx, "_precision", "f64")
    key = (ctx.I, ctx.J, ctx.K, str(ctx.T), prec)
    if key not in _cache:
        choice = {"waves": None, "split_q": -1, "split_p": -1, "mfma_q": -1, "mfma_p": -1}
        uq = lambda: ctx.update_q(False)  # noqa: E731
        up = lambda: ctx.update_p(False)  # noqa: E731
        ctx.profile(True)
        try:
            for tk in (N.ADMX_TUNE_SPLIT_Q, N.ADMX_TUNE_SPLIT_P, N.ADMX_TUNE_QSIDE_MFMA, N.ADMX_TUNE_PSIDE_MFMA):
                ctx.set_tuning(tk, -1)
            best = float("inf")
            for w in candidates:
                ctx.set_tuning(N.ADMX_TUNE_TARGET_WAVES, w)
                ms = _time(ctx, (uq, up), reps)
                if ms < best:
                    choice["waves"], best = w, ms
            ctx.set_tuning(N.ADMX_TUNE_TARGET_WAVES, choice["waves"])
            sides = ()
            if ctx.K >= 9 and prec == "f64":
                sides = (("split_q", N.ADMX_TUNE_SPLIT_Q, uq), ("split_p", N.ADMX_TUNE_SPLIT_P, up))
            elif prec == "bf16":
                sides = (("mfma_q", N.ADMX_TUNE_QSIDE_MFMA, uq), ("mfma_p", N.ADMX_TUNE_PSIDE_MFMA, up))
            if sides:
                for side, tkey, call in sides:
                    times = {}
                    for v in (0, 1):
                        ctx.set_tuning(tkey, v)
                        times[v] = _time(ctx, (call,), reps)
                    choice[side] = min(times, key=times.get)
                    ctx.set_tuning(tkey, choice[side])
        finally:
            ctx.profile(False)
        _cache[key] = choice
    choice = _cache[key]
    ctx.set_tuning(N.ADMX_TUNE_TARGET_WAVES, choice["waves"])
    ctx.set_tuning(N.ADMX_TUNE_SPLIT_Q, choice["split_q"])
    ctx.set_tuning(N.ADMX_TUNE_SPLIT_P, choice["split_p"])
    ctx.set_tuning(N.ADMX_TUNE_QSIDE_MFMA, choice["mfma_q"])
    ctx.set_tuning(N.ADMX_TUNE_PSIDE_MFMA, choice["mfma_p"])
    return dict(choice)


def autotune_waves(ctx, candidates=CANDIDATES, reps: int = 2) -> int:
    """autotune(), returning the chosen waves per pass."""
    return autotune(ctx, candidates, reps)["waves"]
