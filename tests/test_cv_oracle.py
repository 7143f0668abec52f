"""The cross-validation holdout of the K sweep, on the CPU oracle.

The reference has no cross-validation, so these functions are the project's
definition (parity against the reference unpinned; documented as such).  They
are pinned here by an independent numpy restatement of the fold hash, the
partition property of the folds, the masked-BED contract against the decoded
genotypes, and closed-form deviance values.
"""
import numpy as np
import pytest

from conftest import make_cohort

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _smix(z):
    with np.errstate(over="ignore"):
        z = (z + np.uint64(0x9E3779B97F4A7C15)) & M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & M64
        return z ^ (z >> np.uint64(31))


def np_folds(I, J, seed, nfolds):
    """fold[i, j] from the published splitmix64 finalizer, in numpy uint64."""
    base = _smix(np.uint64(seed) ^ np.uint64(0x2545F4914F6CDD1D))
    ctr = (np.arange(J, dtype=np.uint64)[None, :] * np.uint64(I) + np.arange(I, dtype=np.uint64)[:, None])
    h = _smix(ctr ^ base) >> np.uint64(32)
    with np.errstate(over="ignore"):
        return ((h * np.uint64(nfolds)) >> np.uint64(32)).astype(np.int64)


def test_fold_hash_matches_numpy_restatement(O):
    I, J = 37, 29
    F = np_folds(I, J, 11, 5)
    got = np.array([[O.fold_of(i, j, I, 11, 5) for j in range(J)] for i in range(I)])
    assert np.array_equal(got, F)


@pytest.mark.parametrize("nfolds", [2, 5, 10])
def test_folds_partition_entries_evenly(nfolds):
    I, J = 400, 500
    F = np_folds(I, J, 3, nfolds)
    assert F.min() == 0 and F.max() == nfolds - 1
    n = I * J
    counts = np.bincount(F.ravel(), minlength=nfolds)
    # binomial(n, 1/nfolds): every fold within 5 sigma of n/nfolds
    sd = np.sqrt(n * (1 / nfolds) * (1 - 1 / nfolds))
    assert np.all(np.abs(counts - n / nfolds) < 5 * sd), counts
    # another seed gives an independent partition: entries differ w.p. 1 - 1/nfolds
    assert (np_folds(I, J, 4, nfolds) != F).mean() > 0.95 * (1 - 1 / nfolds)


def test_mask_fold_sets_exactly_the_fold_missing(O):
    I, J, K = 203, 151, 3
    _, _, bed, pitch = make_cohort(O, I, J, K, miss=0.05, seed=5)
    g0 = O.decode(bed, pitch, I, J)
    F = np_folds(I, J, 9, 5)
    for f in range(5):
        m = O.mask_fold(bed, pitch, I, J, f, 5, 9)
        g = O.decode(m, pitch, I, J)
        assert np.all(g[F == f] == -1)
        assert np.array_equal(g[F != f], g0[F != f])
        # padding bits of the last byte untouched
        assert np.array_equal(m.reshape(J, pitch)[:, -1] & ~np.uint8((1 << (2 * (I % 4))) - 1 if I % 4 else 0),
                              bed.reshape(J, pitch)[:, -1] & ~np.uint8((1 << (2 * (I % 4))) - 1 if I % 4 else 0))


def _dev_numpy(g, s):
    d = np.zeros_like(s)
    gp, gm = g > 0, g < 2
    d[gp] += g[gp] * np.log(g[gp] / (2 * s[gp]))
    d[gm] += (2 - g[gm]) * np.log((2 - g[gm]) / (2 * (1 - s[gm])))
    return 2 * d


def test_heldout_deviance_against_brute_force(O):
    I, J, K = 120, 90, 3
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, miss=0.03, seed=8)
    g = O.decode(bed, pitch, I, J).astype(np.float64)
    F = np_folds(I, J, 2, 5)
    s = Qt.T @ Pt
    for f in range(5):
        dev, cnt = O.heldout_deviance(bed, pitch, Qt, Pt, f, 5, 2)
        sel = (F == f) & (g >= 0)
        assert cnt == sel.sum()
        ref = _dev_numpy(g[sel], s[sel]).sum()
        assert abs(dev - ref) <= 1e-10 * abs(ref)


def test_deviance_closed_forms():
    # a heterozygote at s = 1/2 has zero deviance; a homozygote g=2 at s gives -4 log s
    g = np.array([1.0, 2.0, 0.0])
    s = np.array([0.5, 0.3, 0.3])
    d = _dev_numpy(g, s)
    assert abs(d[0]) < 1e-15
    assert abs(d[1] - (-4 * np.log(0.3))) < 1e-14
    assert abs(d[2] - (-4 * np.log(0.7))) < 1e-14
