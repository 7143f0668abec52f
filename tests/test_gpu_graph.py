"""hipGraph replay of steady-state admixture_qn! iterations: every replayed
iteration must reproduce the eager iteration bit for bit (the kernels are
deterministic), with eager library work and synchronisations between replays,
across buffer-layout changes that force a re-capture."""
import numpy as np
import pytest

from conftest import make_cohort

pytestmark = pytest.mark.gpu


def _pair(oa, O, I, J, K, seed, form):
    Qt, Pt, bed, pitch = make_cohort(O, I, J, K, alpha=0.4, miss=0.02, seed=seed)
    T = np.float32 if form == "ctx_f32" else np.float64  # ctx_f32: FP32 context (FP64 passes on widened q, p)
    out = []
    for graphs in (False, True):
        c = oa.HIPAdmixData(I, J, K, 3, T)
        c.upload_bed(bed.reshape((pitch, J), order="F"))
        if T == np.float64:
            c.set_precision("bf16" if form == "bf16_mfma" else form)
        if form == "bf16_mfma":  # the matrix-core update passes (LDS image copies, barriers) under capture
            c.set_tuning(oa._native.ADMX_TUNE_QSIDE_MFMA, 1)
            c.set_tuning(oa._native.ADMX_TUNE_PSIDE_MFMA, 1)
        c.set_graphs(graphs)
        d = oa.AdmixData(I, J, K, 3, rng=np.random.default_rng(seed + 1), T=T)
        oa.init_em(d, None, 3, d_cu=c, log=None)
        c.set_uv(d.U, d.V)
        c.set_ll(d.ll_prev, d.ll_new)
        out.append(c)
    return out


@pytest.mark.parametrize("form", ["f64", "bf16", "bf16_mfma", "ctx_f32"])
def test_graph_replay_is_bit_identical(oa, O, form):
    print("graph replay", form, flush=True)
    I, J, K = 900, 2500, 4
    eager, graph = _pair(oa, O, I, J, K, 31, form)
    # Q = 3: iterations 4-6 capture one graph per history column, 7-8 replay
    # them, 9 changes the launch shape (graphs dropped), 9-10 run eagerly,
    # 11-13 re-capture and 14-16 replay.
    for it in range(1, 17):
        if it == 9:
            for c in (eager, graph):
                c.set_tuning(oa._native.ADMX_TUNE_TARGET_WAVES, 8192)
        a = eager.qn_iteration(it)
        b = graph.qn_iteration(it)
        assert a == b, (it, a, b)
        # eager work and a synchronisation between replays
        ll = graph.loglikelihood(oa._native.ADMX_X_NEXT2)
        assert ll == eager.loglikelihood(oa._native.ADMX_X_NEXT2)
        graph.synchronize()
    assert np.array_equal(eager.get_x(0), graph.get_x(0))
    U1, V1 = eager.get_uv()
    U2, V2 = graph.get_uv()
    assert np.array_equal(U1, U2) and np.array_equal(V1, V2)
    eager.close()
    graph.close()
