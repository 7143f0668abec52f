"""SKFR AIM pre-selection on the device (src/driver.jl:56-123) against the CPU
oracle: k-means++ labels, cluster labels, AIM lists and iteration counts
bit-exact for the same nsplit; centers, WSS and objective to 1e-12 relative;
the driver writes the filtered PLINK trio and fits on it."""
import os

import numpy as np
import pytest

from test_skfr_oracle import cohort, pack

pytestmark = pytest.mark.gpu


def _ctx(oa, G):
    I, J = G.shape
    bed = pack(G)
    c = oa.HIPAdmixData(I, J, 2, 1)
    c.upload_bed(bed)
    return c, bed


def _close(a, b, rtol=1e-12):
    a, b = np.asarray(a, float), np.asarray(b, float)
    scale = max(1.0, float(np.abs(b).max())) if b.size else 1.0
    assert np.abs(a - b).max() <= rtol * scale, (np.abs(a - b).max(), scale)


@pytest.mark.parametrize("I,J,k,nsplit", [(37, 29, 1, 1), (203, 151, 3, 2), (333, 410, 8, 7), (130, 96, 16, 0)])
def test_kmeanspp_matches_oracle(oa, O, I, J, k, nsplit):
    print("skfr init", I, J, k, nsplit, flush=True)
    G, _ = cohort(I, J, min(k, 4), seed=I + k, miss=0.05)
    c, bed = _ctx(oa, G)
    for seed in (0, 12345):
        got = c.skfr_init(k, seed, nsplit)
        want, _ = O.skfr_init(bed, I, J, k, seed, nsplit)
        np.testing.assert_array_equal(got, want)
    c.close()


@pytest.mark.parametrize("mode", ["global", "local"])
@pytest.mark.parametrize("I,J,k,s,nsplit", [(203, 151, 3, 20, 2), (517, 700, 5, 60, 9), (1001, 333, 12, 40, 0),
                                            (64, 50, 16, 50, 1)])
def test_run_matches_oracle(oa, O, mode, I, J, k, s, nsplit):
    print("skfr run", mode, I, J, k, s, nsplit, flush=True)
    G, _ = cohort(I, J, min(k, 5), seed=3 * I + J, miss=0.03, aims=np.arange(0, J, 4))
    c, bed = _ctx(oa, G)
    start = c.skfr_init(k, 7, nsplit)
    got = c.skfr_run(k, start, s, mode=mode, max_iter=40, nsplit=nsplit)
    want = O.skfr_run(bed, I, J, k, start, s, mode, 40, nsplit)
    np.testing.assert_array_equal(got[0], want[0])  # clusters
    np.testing.assert_array_equal(got[2], want[2])  # aims
    assert got[5] == want[5]  # iterations
    _close(got[1], want[1])  # centers
    _close(got[3], want[3])  # WSS
    assert abs(got[4] - want[4]) <= 1e-12 * abs(want[4])
    c.close()


def test_run_max_iter_and_restart(oa, O):
    print("skfr max_iter", flush=True)
    I, J, k = 400, 300, 4
    G, _ = cohort(I, J, k, seed=21, miss=0.02)
    c, bed = _ctx(oa, G)
    start = np.random.default_rng(0).integers(0, k, size=I).astype(np.int32)
    for it in (1, 2, 3):
        got = c.skfr_run(k, start, 25, max_iter=it, nsplit=3)
        want = O.skfr_run(bed, I, J, k, start, 25, "global", it, 3)
        np.testing.assert_array_equal(got[0], want[0])
        np.testing.assert_array_equal(got[2], want[2])
        assert got[5] == want[5] <= it
    c.close()


def test_larger_cohort_auto_split(oa, O):
    print("skfr larger", flush=True)
    I, J, k, s = 2000, 3000, 6, 300
    G, lab = cohort(I, J, k, seed=5, miss=0.01, aims=np.arange(0, J, 5))
    c, bed = _ctx(oa, G)
    start = c.skfr_init(k, 99, 0)
    got = c.skfr_run(k, start, s, max_iter=10, nsplit=0)
    want = O.skfr_run(bed, I, J, k, start, s, "global", 10, 0)
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_array_equal(got[2], want[2])
    assert got[5] == want[5]
    assert abs(got[4] - want[4]) <= 1e-12 * abs(want[4])
    c.close()


def test_package_entry_points(oa, O):
    print("skfr package", flush=True)
    I, J, k = 300, 240, 3
    G, lab = cohort(I, J, k, seed=8, aims=np.arange(0, J, 3))
    c, bed = _ctx(oa, G)
    X = oa.ImputedSnpMatrix(c, k, rng=4, nsplit=2)
    want0, _ = O.skfr_init(bed, I, J, k, 4, 2)
    np.testing.assert_array_equal(X.clusters, want0)
    cl, cen, aims, wss, obj = oa.sparsekmeans1(X, 30, max_iter=20)
    w = O.skfr_run(bed, I, J, k, want0, 30, "global", 20, 2)
    np.testing.assert_array_equal(cl, w[0])
    np.testing.assert_array_equal(aims, w[2])
    assert cen.shape == (k, J) and wss.shape == (k,)
    r = oa.sparsekmeans_repeat(X, 30, iter=3, max_inner_iter=20)
    assert len(r) == 6 and r[4] == min(r[5]) and r[4] <= obj + 1e-9 * abs(obj)
    cls, aimss = oa.sparsekmeans_path(X, [60, 30, 10], iter=2, max_inner_iter=20)
    assert [len(a) for a in aimss] == [60, 30, 10]
    cl2, _, aims2, _, _ = oa.sparsekmeans2(X, 10, max_iter=20)
    assert 10 <= len(aims2) <= 3 * 10 and len(set(aims2.tolist())) == len(aims2)
    c.close()


def test_argument_errors(oa, O):
    G, _ = cohort(50, 40, 2, seed=1)
    c, _ = _ctx(oa, G)
    start = np.zeros(50, np.int32)
    with pytest.raises(oa.AdmixHIPError):
        c.skfr_run(3, start + 3, 5)  # label out of range
    with pytest.raises(oa.AdmixHIPError):
        c.skfr_run(3, start, 41)  # sparsity > J
    with pytest.raises(oa.AdmixHIPError):
        c.skfr_run(17, start, 5)  # k > 16
    with pytest.raises(oa.AdmixHIPError):
        c.skfr_init(51, 0)  # more clusters than individuals
    with pytest.raises(ValueError):
        c.skfr_run(3, start, 5, mode="other")
    c.close()


def test_run_admixture_with_sparsity(oa, O, tmp_path):
    print("run_admixture sparsity", flush=True)
    I, J, K, s = 240, 900, 3, 150
    rng = np.random.default_rng(2)
    lab = rng.integers(0, K, size=I)
    P = np.tile(rng.uniform(0.1, 0.9, size=J), (K, 1))
    inf = rng.choice(J, 120, replace=False)
    for j in inf:
        P[:, j] = rng.permutation([0.1, 0.5, 0.9])
    G = rng.binomial(2, P[lab]).astype(np.int64)
    G[rng.random((I, J)) < 0.01] = -1
    bed = pack(G)
    stem = str(tmp_path / "cohort")
    oa.write_plink(stem, bed, I, J)
    d, clusters, aims = oa.run_admixture(stem + ".bed", K, rng=np.random.default_rng(3), sparsity=s,
                                         prefix=str(tmp_path / "sel"), admix_n_iter=30, log=None)
    out = str(tmp_path / f"sel_{K}_{s}aims.bed")
    assert os.path.exists(out)
    assert len(aims) == s and len(set(aims.tolist())) == s
    sub = oa.SnpArray(out)
    assert sub.shape == (I, s)
    np.testing.assert_array_equal(np.asarray(sub.data), bed[:, np.sort(aims)])
    with open(out[:-4] + ".bim") as f:
        names = [line.split()[1] for line in f]
    assert names == [f"snp{j}" for j in np.sort(aims)]
    assert d.J == s and np.isfinite(d.ll_new)
    # most selected SNPs are the informative ones
    assert np.isin(aims, inf).mean() > 0.7
    assert clusters.shape == (I,)
