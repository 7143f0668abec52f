"""AdmixData{T} (src/structs.jl:9-215): the algorithm's host-side state.

Layout is the reference's: x is K x (I+J) column-major, q = x[:, :I] and
p = x[:, I:] are views of it (src/structs.jl:120-137), XtX is K x K x n, U and
V are K(I+J) x Q.  Initialisation follows src/structs.jl:105-215 with numpy's
Generator in place of Julia's RNG: x ~ U(0,1), columns of q normalised to
sum 1, log-likelihoods NaN.  The reference also fills XtX/Xtz/U/V with U(0,1)
draws; every use overwrites them first (update_q!/update_p! write XtX/Xtz,
admixture_qn! reads only the U/V columns it has written,
src/algorithms_outer.jl:125-133), so here they start zeroed, which allocates
lazily instead of drawing ~10 K^2 (I + J) random numbers per fit.
"""
from __future__ import annotations

import numpy as np


class AdmixData:
    def __init__(self, I: int, J: int, K: int, Q: int = 3, skipmissing: bool = True, rng=None, T=np.float64):
        rng = np.random.default_rng() if rng is None else rng
        T = np.dtype(T)
        self.I, self.J, self.K, self.Q = int(I), int(J), int(K), int(Q)
        self.T = T
        self.skipmissing = bool(skipmissing)
        self.x = np.asfortranarray(rng.random((K, I + J)).astype(T))
        self.x_next = np.zeros((K, I + J), T, order="F")
        self.x_next2 = np.zeros((K, I + J), T, order="F")
        self.x_tmp = np.zeros((K, I + J), T, order="F")
        q = self.x[:, :I]
        q /= q.sum(axis=0, keepdims=True)
        self.XtX_q = np.zeros((K, K, I), T, order="F")
        self.Xtz_q = np.zeros((K, I), T, order="F")
        self.XtX_p = np.zeros((K, K, J), T, order="F")
        self.Xtz_p = np.zeros((K, J), T, order="F")
        self.V = np.zeros((K * (I + J), Q), T, order="F")
        self.U = np.zeros((K * (I + J), Q), T, order="F")
        self.ll_prev = float("nan")
        self.ll_new = float("nan")

    # views aliasing x, x_next, x_next2, x_tmp (src/structs.jl:120-137)
    @property
    def q(self):
        return self.x[:, : self.I]

    @property
    def p(self):
        return self.x[:, self.I:]

    @property
    def q_next(self):
        return self.x_next[:, : self.I]

    @property
    def p_next(self):
        return self.x_next[:, self.I:]

    @property
    def q_next2(self):
        return self.x_next2[:, : self.I]

    @property
    def p_next2(self):
        return self.x_next2[:, self.I:]

    @property
    def q_tmp(self):
        return self.x_tmp[:, : self.I]

    @property
    def p_tmp(self):
        return self.x_tmp[:, self.I:]

    @property
    def x_flat(self):
        return self.x.reshape(-1, order="F")
