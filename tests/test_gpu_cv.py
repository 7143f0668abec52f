"""Device cross-validation holdout and the K sweep (no reference counterpart;
checked against the CPU oracle's restatement, bit-exact for the masked BED and
to 1e-12 relative for the held-out deviance, and end to end by recovering the
true K of a synthetic cohort)."""
import numpy as np
import pytest

from conftest import make_cohort, random_state

pytestmark = pytest.mark.gpu


def _src(oa, O, I, J, K, seed, miss=0.02, alpha=0.5):
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, alpha=alpha, miss=miss, seed=seed)
    c = oa.HIPAdmixData(I, J, 2, 3)
    c.upload_bed(bed.reshape((pitch, J), order="F"))
    return c, Qt, Pt, bed, pitch


@pytest.mark.parametrize("I,J", [(203, 151), (1000, 777)])
def test_bed_from_masks_exactly_the_fold(oa, O, I, J):
    print("bed_from", I, J, flush=True)
    src, _, _, bed, pitch = _src(oa, O, I, J, 3, seed=4)
    dst = oa.HIPAdmixData(I, J, 4, 3)
    dst.bed_from(src, -1)
    assert np.array_equal(dst.download_bed().ravel(order="F"), bed)
    for f in (0, 3):
        dst.bed_from(src, f, 5, 17)
        ref = O.mask_fold(bed, pitch, I, J, f, 5, 17)
        assert np.array_equal(dst.download_bed().ravel(order="F"), ref), f
    # the masked context still computes: counts agree with the oracle's
    snm, salt, inm = dst.counts()
    rs, ra, ri = O.counts(ref, pitch, I, J)
    assert np.array_equal(snm, rs) and np.array_equal(salt, ra) and np.array_equal(inm, ri)
    dst.close()
    src.close()


def test_heldout_deviance_matches_oracle(oa, O):
    print("heldout deviance", flush=True)
    I, J, K = 700, 1300, 4
    src, _, _, bed, pitch = _src(oa, O, I, J, 3, seed=6)
    fit = oa.HIPAdmixData(I, J, K, 3)
    q, p = random_state(I, J, K, seed=2)
    fit.set_x(0, np.asfortranarray(np.hstack([q, p])))
    for f in range(5):
        dev, cnt = fit.heldout_deviance(src, f, 5, 23)
        rdev, rcnt = O.heldout_deviance(bed, pitch, q, p, f, 5, 23)
        assert cnt == rcnt
        assert abs(dev - rdev) <= 1e-12 * abs(rdev), (f, dev, rdev)
    fit.close()
    src.close()


def test_cv_argument_errors(oa, O):
    src, *_ = _src(oa, O, 100, 60, 2, seed=1)
    other = oa.HIPAdmixData(101, 60, 3, 3)
    with pytest.raises(oa.AdmixHIPError):
        other.bed_from(src, 0, 5, 1)
    same = oa.HIPAdmixData(100, 60, 3, 3)
    with pytest.raises(oa.AdmixHIPError):
        same.bed_from(src, 5, 5, 1)
    with pytest.raises(oa.AdmixHIPError):
        same.heldout_deviance(src, -1, 5, 1)
    for c in (other, same, src):
        c.close()


def test_cv_sweep_recovers_true_k(oa, O):
    print("cv sweep", flush=True)
    I, J, Ktrue = 800, 6000, 3
    src, *_ = _src(oa, O, I, J, Ktrue, seed=12, alpha=0.5)
    res = oa.cv_sweep(src, [2, 3, 4, 5], nfolds=5, fold_seed=3, max_iter=300, autotune=False)
    errs = {r["K"]: r["cv_error"] for r in res}
    print(errs, flush=True)
    assert oa.best_k(res) == Ktrue
    assert errs[2] > errs[3]
    for r in res:
        assert r["converged_full"] and sum(r["fold_entries"]) > 0.97 * I * J
    src.close()


def test_folds_from_full_fit_match_random_starts(oa, O):
    """Fold fits started from the full fit's estimates reach the same CV
    errors (to the convergence tolerance) as fold fits from their own random
    starts, in fewer iterations."""
    print("cv fold starts", flush=True)
    I, J, Ktrue = 600, 4000, 3
    src, *_ = _src(oa, O, I, J, Ktrue, seed=21, alpha=0.5)
    warm = oa.cv_sweep(src, [2, 3, 4], nfolds=4, fold_seed=5, max_iter=300, autotune=False, fold_start="full")
    cold = oa.cv_sweep(src, [2, 3, 4], nfolds=4, fold_seed=5, max_iter=300, autotune=False, fold_start="random")
    assert oa.best_k(warm) == oa.best_k(cold) == Ktrue
    for w, c in zip(warm, cold):
        print(w["K"], w["cv_error"], c["cv_error"], w["iterations_folds"], c["iterations_folds"], flush=True)
        assert sum(w["iterations_folds"]) < sum(c["iterations_folds"])
        if w["K"] == Ktrue:  # a well-identified K: both reach the same optimum
            assert abs(w["cv_error"] - c["cv_error"]) <= 1e-4 * c["cv_error"]
    with pytest.raises(ValueError):
        oa.cv_sweep(src, [2], fold_start="other")
    src.close()


def test_autotune_picks_a_candidate(oa, O):
    src, *_ = _src(oa, O, 900, 3000, 3, seed=2)
    c = oa.HIPAdmixData(900, 3000, 3, 3)
    c.bed_from(src, -1)
    q, p = random_state(900, 3000, 3, seed=1)
    c.set_x(0, np.asfortranarray(np.hstack([q, p])))
    w = oa.autotune_waves(c, candidates=(4096, 16384, 65536), reps=1)
    assert w in (4096, 16384, 65536)
    c.close()
    src.close()


def test_concurrent_sweep_equals_serial(oa, O):
    """cv_sweep(workers=4): the K values run at once, each context on its own
    stream from its own host thread; every K's CV error, fold deviances, full
    LL and iteration counts are those of the serial sweep, bit for bit (the
    autotuner runs alone first, so the chunking is the same)."""
    print("cv sweep concurrent", flush=True)
    I, J, Ktrue = 700, 5000, 3
    src, *_ = _src(oa, O, I, J, Ktrue, seed=14, alpha=0.5)
    kw = dict(nfolds=3, fold_seed=2, max_iter=200)
    serial = oa.cv_sweep(src, [2, 3, 4, 5], **kw)
    conc = oa.cv_sweep(src, [2, 3, 4, 5], workers=4, **kw)
    for a, b in zip(serial, conc):
        print(a["K"], a["cv_error"], b["cv_error"], a["iterations_full"], b["iterations_full"], flush=True)
        assert a["K"] == b["K"]
        for key in ("cv_error", "ll_full", "iterations_full", "iterations_folds", "fold_deviance", "waves"):
            assert a[key] == b[key], (a["K"], key, a[key], b[key])
    src.close()
