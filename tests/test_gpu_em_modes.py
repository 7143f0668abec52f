"""em_q! / em_p! with the reference's three modes (src/algorithms_inner.jl:54-141)
through the host mirror and the C ABI (admx_em_q_mode / admx_em_p_mode),
against the oracle's restatement of the same modes, on distinct states in x,
x_next and x_next2 so that a wrong source or destination shows:

  em_q!  :base, :fallback  q_next  from (q, p)
         :fallback2        q_next2 from (q_next, p_next)
  em_p!  :base             p_next  from (q, p)
         :fallback         p_next  from (q_next, p)
         :fallback2        p_next2 from (q_next2, p_next)

FP64 passes 1e-12 (max-abs; q_next, p_next lie in [0, 1]); the BF16 form's
matrix-core sums 2e-5.  Every other block of the state is left as it was."""
import numpy as np
import pytest

from conftest import make_cohort

pytestmark = pytest.mark.gpu


def _state(I, J, K, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((K, I + J)) + 0.05
    x[:, :I] /= x[:, :I].sum(0)
    x[:, I:] = rng.uniform(0.02, 0.98, (K, J))
    return np.asfortranarray(x)


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("bf16", 2e-5)])
@pytest.mark.parametrize("side", ["q", "p"])
@pytest.mark.parametrize("mode", ["base", "fallback", "fallback2"])
def test_em_modes_match_oracle(oa, O, side, mode, prec, tol):
    print("em mode", side, mode, prec, flush=True)
    I, J, K = 450, 700, 4
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, miss=0.02, seed=31)
    d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(1))
    d.x[...], d.x_next[...], d.x_next2[...] = _state(I, J, K, 2), _state(I, J, K, 3), _state(I, J, K, 4)
    before = [a.copy() for a in (d.x, d.x_next, d.x_next2)]
    ref = O.OracleAdmix(bed, pitch, I, J, K, 3, d.x)
    ref.x_next[...], ref.x_next2[...] = d.x_next, d.x_next2
    c = oa.HIPAdmixData(I, J, K, 3)
    c.upload_bed(bed.reshape((pitch, J), order="F"))
    c.set_precision(prec)
    if side == "q":
        oa.em_q(d, None, mode=mode, d_cu=c)
        ref.em_q(mode)
    else:
        oa.em_p(d, None, mode=mode, d_cu=c)
        ref.em_p(mode)
    c.close()
    dst = 2 if mode == "fallback2" else 1
    cols = slice(0, I) if side == "q" else slice(I, I + J)
    got = (d.x, d.x_next, d.x_next2)[dst][:, cols]
    want = (ref.x, ref.x_next, ref.x_next2)[dst][:, cols]
    err = np.abs(got - want).max()
    print(f"  max|d| {err:.2e}", flush=True)
    assert err <= tol
    # nothing else of the host state moved
    for n, (a, b) in enumerate(zip((d.x, d.x_next, d.x_next2), before)):
        if n == dst:
            other = slice(I, I + J) if side == "q" else slice(0, I)
            assert np.array_equal(a[:, other], b[:, other])
        else:
            assert np.array_equal(a, b)


def test_em_mode_names(oa, O):
    """Modes are the reference's symbols (with or without the colon); others raise."""
    I, J, K = 40, 50, 2
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, seed=2)
    c = oa.HIPAdmixData(I, J, K, 3)
    c.upload_bed(bed.reshape((pitch, J), order="F"))
    d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(0))
    oa.em_q(d, None, mode=":fallback", d_cu=c)
    with pytest.raises(ValueError):
        oa.em_p(d, None, mode="fallback3", d_cu=c)
    c.close()
