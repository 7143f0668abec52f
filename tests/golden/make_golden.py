"""Writes tests/golden/tiny_f64.npz: a small cohort and the oracle's outputs on it.

The reference's own test data (EUR_subset.bed from SnpArrays, test/runtests.jl)
is not available here and the reference is Julia, which does not run here;
these vectors come from the CPU oracle after it passed the checks in
tests/test_oracle_pins.py, and freeze its outputs so that later changes to
either the oracle or the HIP path are caught against fixed numbers.

Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle as O  # noqa: E402


def main():
    I, J, K, Q = 37, 53, 3, 3
    rng = np.random.default_rng(2024)
    Qt = np.asfortranarray(rng.dirichlet(0.5 * np.ones(K), size=I).T)
    Pt = np.asfortranarray(rng.uniform(0.05, 0.95, size=(K, J)))
    pitch = (I + 3) // 4
    bed = O.gen_bed(Qt, Pt, 99, 0.03, pitch)
    x0 = rng.random((K, I + J))
    x0[:, :I] /= x0[:, :I].sum(0)
    x0 = np.asfortranarray(x0)
    U0 = np.asfortranarray(rng.random((K * (I + J), Q)))
    V0 = np.asfortranarray(rng.random((K * (I + J), Q)))
    q = np.asfortranarray(x0[:, :I])
    p = np.asfortranarray(x0[:, I:])
    out = dict(I=I, J=J, K=K, Q=Q, bed=bed, Qt=Qt, Pt=Pt, x0=x0, U0=U0, V0=V0)
    out["decoded"] = O.decode(bed, pitch, I, J)
    out["snp_nm"], out["snp_alt"], out["ind_nm"] = O.counts(bed, pitch, I, J)
    out["ll"] = np.array(O.loglikelihood(bed, pitch, q, p))
    XtX, Xtz = np.zeros((K, K, I), order="F"), np.zeros((K, I), order="F")
    O.update_q_loop(XtX, Xtz, bed, pitch, q, p)
    out["XtX_q"], out["Xtz_q"] = XtX, Xtz
    XtX, Xtz = np.zeros((K, K, J), order="F"), np.zeros((K, J), order="F")
    O.update_p_loop(XtX, Xtz, bed, pitch, q, p)
    out["XtX_p"], out["Xtz_p"] = XtX, Xtz
    st = O.OracleAdmix(bed, pitch, I, J, K, Q, x0, U0, V0)
    out["em_lls"] = st.init_em(3)
    out["x_after_em"] = st.x.copy(order="F")
    its = [st.qn_iteration(it) for it in (1, 2, 3, 4)]
    out["qn_lls"] = np.array([[r["prev"], r["basic"], r["qn"], r["new"]] for r in its])
    out["x_after_qn"] = st.x.copy(order="F")
    np.savez_compressed(os.path.join(HERE, "tiny_f64.npz"), **out)
    print("wrote tiny_f64.npz", {k: np.shape(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
