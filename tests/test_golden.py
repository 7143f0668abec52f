"""Frozen vectors (tests/golden/tiny_f64.npz, written by make_golden.py): the
oracle must keep reproducing them (CPU), and the HIP path must match them
through the C ABI (GPU)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tiny_f64.npz")


@pytest.fixture(scope="module")
def G():
    return dict(np.load(GOLD, allow_pickle=False))


def test_oracle_reproduces_golden(O, G):
    I, J, K, Q = int(G["I"]), int(G["J"]), int(G["K"]), int(G["Q"])
    pitch = (I + 3) // 4
    bed = O.gen_bed(G["Qt"], G["Pt"], 99, 0.03, pitch)
    assert np.array_equal(bed, G["bed"])
    assert np.array_equal(O.decode(bed, pitch, I, J), G["decoded"])
    q, p = np.asfortranarray(G["x0"][:, :I]), np.asfortranarray(G["x0"][:, I:])
    assert O.loglikelihood(bed, pitch, q, p) == float(G["ll"])
    XtX, Xtz = np.zeros((K, K, I), order="F"), np.zeros((K, I), order="F")
    O.update_q_loop(XtX, Xtz, bed, pitch, q, p)
    assert np.array_equal(XtX, G["XtX_q"]) and np.array_equal(Xtz, G["Xtz_q"])
    st = O.OracleAdmix(bed, pitch, I, J, K, Q, G["x0"], G["U0"], G["V0"])
    assert np.array_equal(st.init_em(3), G["em_lls"])
    its = np.array([[r["prev"], r["basic"], r["qn"], r["new"]] for r in (st.qn_iteration(i) for i in (1, 2, 3, 4))])
    assert np.array_equal(its, G["qn_lls"])
    assert np.array_equal(st.x, G["x_after_qn"])


@pytest.mark.gpu
def test_hip_matches_golden(oa, G):
    I, J, K, Q = int(G["I"]), int(G["J"]), int(G["K"]), int(G["Q"])
    print("golden tiny", I, J, K, flush=True)
    c = oa.HIPAdmixData(I, J, K, Q)
    c.upload_bed(G["bed"])
    assert np.array_equal(c.decode(), G["decoded"])
    snm, salt, inm = c.counts()
    assert np.array_equal(snm, G["snp_nm"]) and np.array_equal(salt, G["snp_alt"]) and np.array_equal(inm, G["ind_nm"])
    q, p = np.asfortranarray(G["x0"][:, :I]), np.asfortranarray(G["x0"][:, I:])
    ll = c.loglikelihood_loop(q, p, 0, I, 0, J)
    assert abs(ll - float(G["ll"])) <= 1e-10 * abs(float(G["ll"]))
    XtX, Xtz = np.zeros((K, K, I), order="F"), np.zeros((K, I), order="F")
    c.update_q_loop(XtX, Xtz, q, p, 0, I, 0, J)
    assert np.allclose(XtX, G["XtX_q"], rtol=1e-9, atol=1e-9 * np.abs(G["XtX_q"]).max())
    assert np.allclose(Xtz, G["Xtz_q"], rtol=1e-9, atol=1e-9 * np.abs(G["Xtz_q"]).max())
    XtX, Xtz = np.zeros((K, K, J), order="F"), np.zeros((K, J), order="F")
    c.update_p_loop(XtX, Xtz, q, p, 0, I, 0, J)
    assert np.allclose(XtX, G["XtX_p"], rtol=1e-9, atol=1e-9 * np.abs(G["XtX_p"]).max())
    assert np.allclose(Xtz, G["Xtz_p"], rtol=1e-9, atol=1e-9 * np.abs(G["Xtz_p"]).max())
    c.set_x(0, np.asfortranarray(G["x0"]))
    c.set_uv(G["U0"], G["V0"])
    lls = c.init_em(3)
    assert np.allclose(lls, G["em_lls"], rtol=1e-10, atol=0)
    for it in (1, 2, 3, 4):
        r = c.qn_iteration(it)
        ref = G["qn_lls"][it - 1]
        got = np.array([r["prev"], r["basic"], r["qn"], r["new"]])
        assert np.allclose(got, ref, rtol=1e-10, atol=0), (it, got, ref)
    assert np.abs(c.get_x(0) - G["x_after_qn"]).max() < 1e-8
    c.close()
