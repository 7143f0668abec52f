import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libadmixhip on the device)")


@pytest.fixture(scope="session")
def oa():
    import __graft_entry__ as g

    return g.load_package()


@pytest.fixture(scope="session")
def O():
    import pyoracle

    pyoracle.lib()
    return pyoracle


def make_cohort(O, I, J, K, alpha=0.5, miss=0.01, seed=1, p_lo=0.05, p_hi=0.95, beta=False):
    """Synthetic cohort: truth Q/P and the packed BED from the oracle's generator."""
    rng = np.random.default_rng(seed)
    Qt = np.asfortranarray(rng.dirichlet(alpha * np.ones(K), size=I).T)
    if beta:
        Pt = np.asfortranarray(np.clip(rng.beta(0.5, 0.5, size=(K, J)), p_lo, p_hi))
    else:
        Pt = np.asfortranarray(rng.uniform(p_lo, p_hi, size=(K, J)))
    pitch = (I + 3) // 4
    bed = O.gen_bed(Qt, Pt, seed, miss, pitch)
    return Qt, Pt, bed, pitch


def random_state(I, J, K, seed=0, T=np.float64):
    """A feasible (q, p) pair: q columns on the simplex, p in (0.02, 0.98)."""
    rng = np.random.default_rng(seed)
    q = rng.random((K, I)) + 0.05
    q /= q.sum(0)
    p = rng.uniform(0.02, 0.98, size=(K, J))
    return np.asfortranarray(q.astype(T)), np.asfortranarray(p.astype(T))
