"""Pins of the SKFR oracle (oracle/skfr_oracle.c): sparse k-means with feature
ranking as the reference's driver calls it (src/driver.jl:56-123, through
SparseKmeansFeatureRanking.jl, a dependency absent from the reference tree).

Each check is computed independently of the oracle's formulas, on a dense
standardized genotype matrix: Lloyd's k-means when every SNP is kept,
brute-force between-cluster sums of squares for the ranking, sparse-center
distances, planted ancestry-informative SNPs, allele-flip invariance and
dense k-means++ draws.
"""
import numpy as np
import pytest

CODE_OF_G = np.array([0b00, 0b10, 0b11], np.uint8)  # g = 0, 1, 2; missing = 0b01


def pack(G):
    """BED payload (ceil(I/4) x J, Fortran) of genotypes G (I x J, -1 missing)."""
    I, J = G.shape
    codes = np.where(G < 0, 0b01, CODE_OF_G[np.clip(G, 0, 2)]).astype(np.uint8)
    nb = (I + 3) // 4
    out = np.zeros((nb, J), np.uint8)
    for t in range(4):
        part = codes[t::4]
        out[: part.shape[0]] |= part << (2 * t)
    return np.asfortranarray(out)


def cohort(I, J, k, seed, miss=0.02, aims=None, sep=0.35):
    """Genotypes of k populations; with `aims`, only those SNPs differ in frequency."""
    rng = np.random.default_rng(seed)
    lab = rng.integers(0, k, size=I)
    base = rng.uniform(0.1, 0.9, size=J)
    P = np.tile(base, (k, 1))
    cols = np.arange(J) if aims is None else np.asarray(aims)
    P[:, cols] = np.clip(base[cols] + rng.uniform(-sep, sep, size=(k, cols.size)), 0.02, 0.98)
    G = rng.binomial(2, P[lab]).astype(np.int64)
    G[rng.random((I, J)) < miss] = -1
    return G, lab


def dense_x(G):
    obs = G >= 0
    nm = obs.sum(0)
    mu = np.where(nm > 0, np.where(obs, G, 0).sum(0) / np.maximum(nm, 1), 0.0)
    v = mu * (1 - mu / 2)
    sd = np.sqrt(v)
    X = np.where(obs & (v > 0), (G - mu) / np.where(sd > 0, sd, 1), 0.0)
    return X, mu, v


def lloyd(X, lab, k, max_iter):
    lab = lab.copy()
    for it in range(1, max_iter + 1):
        C = np.zeros((k, X.shape[1]))
        n = np.bincount(lab, minlength=k)
        for c in range(k):
            if n[c]:
                C[c] = X[lab == c].mean(0)
        D = ((X[:, None, :] - C[None]) ** 2).sum(2)
        D[:, n == 0] = np.inf
        new = D.argmin(1)
        changed = (new != lab).sum()
        lab = new
        if changed == 0:
            break
    return lab, it, D[np.arange(len(lab)), lab].sum()


def bss(X, lab, k):
    n = np.bincount(lab, minlength=k)
    s = np.zeros(X.shape[1])
    per = np.zeros((k, X.shape[1]))
    for c in range(k):
        if n[c]:
            per[c] = n[c] * X[lab == c].mean(0) ** 2
    s = per.sum(0)
    return s, per, n


def rank(score):
    return np.lexsort((np.arange(score.size), -score))


def test_stats_are_observed_means(O):
    G, _ = cohort(37, 21, 3, 1, miss=0.2)
    G[:, 4] = -1  # all missing
    G[:, 5] = 2  # monomorphic
    mu, v = O.skfr_stats(pack(G), 37, 21)
    X, mu_d, v_d = dense_x(G)
    np.testing.assert_allclose(mu, mu_d, rtol=0, atol=1e-15)
    np.testing.assert_allclose(v, v_d, rtol=0, atol=1e-15)
    assert mu[4] == 0 and v[4] == 0 and mu[5] == 2 and v[5] == 0


@pytest.mark.parametrize("nsplit", [1, 3])
def test_full_sparsity_is_lloyds_kmeans(O, nsplit):
    I, J, k = 90, 40, 3
    G, _ = cohort(I, J, k, 2)
    rng = np.random.default_rng(0)
    start = rng.integers(0, k, size=I).astype(np.int32)
    cl, cen, aims, wss, obj, iters = O.skfr_run(pack(G), I, J, k, start, J, "global", 30, nsplit)
    X, _, _ = dense_x(G)
    lab_d, it_d, obj_d = lloyd(X, start, k, 30)
    np.testing.assert_array_equal(cl, lab_d)
    assert iters == it_d
    assert obj == pytest.approx(obj_d, rel=1e-10)
    assert sorted(aims) == list(range(J))
    # centers are the cluster means of the clusters the last assignment used
    # (at convergence: of the final labels)
    for c in range(k):
        np.testing.assert_allclose(cen[c], X[cl == c].mean(0), rtol=0, atol=1e-12)
        assert wss[c] == pytest.approx(((X[cl == c] - cen[c]) ** 2).sum(), rel=1e-10)


def test_one_iteration_ranks_by_between_cluster_ss(O):
    I, J, k, s = 120, 60, 4, 9
    G, lab = cohort(I, J, k, 3, aims=np.arange(0, 60, 3))
    rng = np.random.default_rng(1)
    start = np.where(rng.random(I) < 0.8, lab, rng.integers(0, k, size=I)).astype(np.int32)
    cl, cen, aims, wss, obj, iters = O.skfr_run(pack(G), I, J, k, start, s, "global", 1, 2)
    X, _, _ = dense_x(G)
    score, _, n = bss(X, start, k)
    want = rank(score)[:s]
    np.testing.assert_array_equal(aims, want)
    # sparse centers: cluster means on the selected SNPs, 0 elsewhere
    C = np.zeros((k, J))
    for c in range(k):
        C[c, want] = X[start == c][:, want].mean(0)
    np.testing.assert_allclose(cen, C, atol=1e-12)
    D = ((X[:, None, :] - C[None]) ** 2).sum(2)
    np.testing.assert_array_equal(cl, D.argmin(1))
    assert obj == pytest.approx(D.min(1).sum(), rel=1e-10)


def test_local_ranking_per_cluster(O):
    I, J, k, s = 100, 50, 3, 5
    G, lab = cohort(I, J, k, 4, aims=np.arange(0, 50, 2))
    start = lab.astype(np.int32)
    cl, cen, aims, wss, obj, iters = O.skfr_run(pack(G), I, J, k, start, s, "local", 1, 1)
    X, _, _ = dense_x(G)
    _, per, n = bss(X, start, k)
    sel = np.zeros((k, J), bool)
    for c in range(k):
        sel[c, rank(per[c])[:s]] = True
    union = sel.any(0)
    ukey = np.where(union, np.where(sel, per, -1).max(0), -1.0)
    np.testing.assert_array_equal(aims, rank(ukey)[: union.sum()])
    C = np.zeros((k, J))
    for c in range(k):
        C[c, sel[c]] = X[start == c][:, sel[c]].mean(0)
    np.testing.assert_allclose(cen, C, atol=1e-12)
    D = ((X[:, None, :] - C[None]) ** 2).sum(2)
    np.testing.assert_array_equal(cl, D.argmin(1))


def test_planted_aims_are_recovered(O):
    I, J, k = 400, 200, 3
    rng = np.random.default_rng(5)
    planted = np.sort(rng.choice(J, 12, replace=False))
    lab = rng.integers(0, k, size=I)
    P = np.tile(rng.uniform(0.1, 0.9, size=J), (k, 1))
    for j in planted:  # well-separated frequencies on the planted SNPs only
        P[:, j] = rng.permutation([0.1, 0.5, 0.9])
    G = rng.binomial(2, P[lab]).astype(np.int64)
    G[rng.random((I, J)) < 0.02] = -1
    cl, cen, aims, wss, obj, iters = O.skfr_run(pack(G), I, J, k, lab.astype(np.int32), 12, "global", 20, 4)
    assert set(aims.tolist()) == set(planted.tolist())
    # clusters equal the populations up to relabelling (almost all individuals)
    agree = max((cl == lab).mean(), 0)
    assert agree > 0.95


def test_allele_flip_invariance(O):
    I, J, k, s = 150, 80, 3, 20
    G, lab = cohort(I, J, k, 6)
    start = np.random.default_rng(2).integers(0, k, size=I).astype(np.int32)
    a = O.skfr_run(pack(G), I, J, k, start, s, "global", 25, 2)
    Gf = np.where(G < 0, -1, 2 - G)
    b = O.skfr_run(pack(Gf), I, J, k, start, s, "global", 25, 2)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[2], b[2])
    assert a[4] == pytest.approx(b[4], rel=1e-9)
    np.testing.assert_allclose(a[1], -b[1], atol=1e-9)


def test_kmeanspp_draws(O):
    I, J, k, seed = 70, 30, 5, 11
    G, _ = cohort(I, J, k, 7)
    cl, centers = O.skfr_init(pack(G), I, J, k, seed, 2)
    X, _, _ = dense_x(G)
    first = min(int(O.skfr_u(seed, 0) * I), I - 1)
    assert centers[0] == first
    D2 = ((X - X[first]) ** 2).sum(1)
    for c in range(1, k):
        target = O.skfr_u(seed, c) * D2.sum()
        want = int(np.searchsorted(np.cumsum(D2), target, side="right"))
        assert centers[c] == want
        D2 = np.minimum(D2, ((X - X[centers[c]]) ** 2).sum(1))
    Dc = np.stack([((X - X[ci]) ** 2).sum(1) for ci in centers], 1)
    np.testing.assert_array_equal(cl, Dc.argmin(1))
    assert len(set(centers.tolist())) == k


def test_empty_cluster_is_never_chosen(O):
    I, J, k = 60, 20, 4
    G, lab = cohort(I, J, 3, 8)
    start = lab.astype(np.int32)  # cluster 3 empty
    cl, cen, aims, wss, obj, iters = O.skfr_run(pack(G), I, J, k, start, 10, "global", 10, 1)
    assert not (cl == 3).any()
    assert np.all(cen[3] == 0) and wss[3] == 0


def test_uninformative_snps_rank_last(O):
    I, J, k = 80, 16, 2
    G, lab = cohort(I, J, k, 9)
    G[:, 3] = 1  # monomorphic het: v > 0 but identical in every cluster
    G[:, 7] = 0  # v = 0
    G[:, 11] = -1  # all missing
    cl, cen, aims, *_ = O.skfr_run(pack(G), I, J, k, lab.astype(np.int32), J, "global", 1, 1)
    assert set(aims[-3:].tolist()) == {3, 7, 11}
    assert np.all(cen[:, [3, 7, 11]] == 0)


def test_nsplit_changes_only_rounding(O):
    I, J, k = 200, 120, 3
    G, lab = cohort(I, J, k, 10, aims=np.arange(40))
    start = np.random.default_rng(3).integers(0, k, size=I).astype(np.int32)
    runs = [O.skfr_run(pack(G), I, J, k, start, 30, "global", 30, ns) for ns in (1, 5, 16)]
    for r in runs[1:]:
        np.testing.assert_array_equal(r[0], runs[0][0])
        np.testing.assert_array_equal(r[2], runs[0][2])
        assert r[4] == pytest.approx(runs[0][4], rel=1e-12)


def test_flat_payload_and_shape_checks(O):
    """A flat payload (as the oracle's generator returns it) is read with
    pitch = size / J; payloads that cannot hold I x J are refused."""
    G, lab = cohort(50, 30, 2, 12)
    bed = pack(G)
    start = lab.astype(np.int32)
    a = O.skfr_run(bed, 50, 30, 2, start, 10, "global", 5, 1)
    b = O.skfr_run(bed.ravel(order="F"), 50, 30, 2, start, 10, "global", 5, 1)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[2], b[2])
    with pytest.raises(ValueError):
        O.skfr_run(bed.ravel(order="F")[:-1], 50, 30, 2, start, 10)
    with pytest.raises(ValueError):
        O.skfr_init(bed[:5], 50, 30, 2, 0, 1)
    with pytest.raises(ValueError):
        O.skfr_run(bed, 50, 30, 2, start + 2, 10)
