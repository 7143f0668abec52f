"""The drop-in workload at full size (I = 10,000, J = 100,000, K = 5, Dirichlet
0.3, 2% missing; 1e9 genotypes): decode and counts bit-exact against the CPU
oracle, the log-likelihood of a state in FP64 and FP32 against the oracle's
loops.  The per-iteration log-likelihoods of this workload in both precisions
are checked in test_gpu_configs.py."""
import numpy as np
import pytest

from conftest import random_state

pytestmark = pytest.mark.gpu

I, J, K = 10_000, 100_000, 5


@pytest.fixture(scope="module")
def cohort(oa):
    from openadmixture_jl_amd import sim

    Qt, Pt = sim.draw_truth(I, J, K, 0.3, "uniform", 0.05, 0.95, seed=2)
    c = oa.HIPAdmixData(I, J, K, 3)
    c.generate_bed(Qt, Pt, 2, 0.02)
    bed = np.ascontiguousarray(c.download_bed().ravel(order="F"))
    yield c, bed
    c.close()


def test_c2_decode_counts_bit_exact(oa, O, cohort):
    print("c2 decode/counts", flush=True)
    c, bed = cohort
    pitch = (I + 3) // 4
    snm, salt, inm = c.counts()
    rs, ra, ri = O.counts(bed, pitch, I, J)
    assert np.array_equal(snm, rs) and np.array_equal(salt, ra) and np.array_equal(inm, ri)
    g = c.decode()
    assert np.array_equal(g, O.decode(bed, pitch, I, J))


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_c2_loglikelihood(oa, O, cohort, T):
    print("c2 LL", np.dtype(T).name, flush=True)
    c, bed = cohort
    pitch = (I + 3) // 4
    q, p = random_state(I, J, K, 3, T)
    if T == np.float64:
        got = c.loglikelihood_loop(q, p, 0, I, 0, J)
    else:
        c32 = oa.HIPAdmixData(I, J, K, 3, T)
        c32.upload_bed(bed.reshape((pitch, J), order="F"))
        got = c32.loglikelihood_loop(q, p, 0, I, 0, J)
        c32.close()
    ref = O.loglikelihood(bed, pitch, q, p)
    tol = 1e-10 if T == np.float64 else 1e-5
    assert abs(got - ref) <= tol * abs(ref), (got, ref)
