"""filter_plink, the SnpArrays.filter step of _filter_SKFR (src/driver.jl:97,
120): SNP columns and individual rows of a PLINK trio, checked against the
decoded genotypes (host file I/O, no GPU)."""
import numpy as np

from test_skfr_oracle import pack


def _decode(data, I):
    codes = np.stack([(data >> (2 * t)) & 3 for t in range(4)], 1).reshape(-1, data.shape[1])
    return codes[:I]


def test_filter_columns_and_rows(oa, tmp_path):
    rng = np.random.default_rng(0)
    I, J = 23, 17
    G = rng.integers(-1, 3, size=(I, J))
    stem = str(tmp_path / "src")
    oa.write_plink(stem, pack(G), I, J)
    cols = np.array([1, 4, 5, 16])
    out = oa.filter_plink(stem + ".bed", None, cols, str(tmp_path / "c"))
    s = oa.SnpArray(out)
    assert s.shape == (I, 4)
    np.testing.assert_array_equal(np.asarray(s.data), pack(G[:, cols]))
    rows = rng.random(I) < 0.6
    out = oa.filter_plink(stem + ".bed", rows, cols, str(tmp_path / "rc"))
    s = oa.SnpArray(out)
    assert s.shape == (rows.sum(), 4)
    np.testing.assert_array_equal(_decode(np.asarray(s.data), rows.sum()), _decode(pack(G[rows][:, cols]), rows.sum()))
    fam = open(str(tmp_path / "rc.fam")).read().split("\n")
    assert [line.split()[1] for line in fam if line] == [f"ind{i}" for i in np.flatnonzero(rows)]
    bim = open(str(tmp_path / "rc.bim")).read().split("\n")
    assert [line.split()[1] for line in bim if line] == [f"snp{j}" for j in cols]
