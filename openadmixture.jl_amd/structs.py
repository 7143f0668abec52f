"""AdmixData{T} (src/structs.jl:9-215): the algorithm's host-side state.

Layout is the reference's: x is K x (I+J) column-major, q = x[:, :I] and
p = x[:, I:] are views of it (src/structs.jl:120-137), XtX is K x K x n, U and
V are K(I+J) x Q.  Initialisation follows src/structs.jl:105-215 with numpy's
Generator in place of Julia's RNG: x ~ U(0,1), columns of q normalised to
sum 1, XtX/Xtz/U/V filled with U(0,1) draws, log-likelihoods NaN.
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
        self.XtX_q = np.asfortranarray(rng.random((K, K, I)).astype(T))
        self.Xtz_q = np.asfortranarray(rng.random((K, I)).astype(T))
        self.XtX_p = np.asfortranarray(rng.random((K, K, J)).astype(T))
        self.Xtz_p = np.asfortranarray(rng.random((K, J)).astype(T))
        self.V = np.asfortranarray(rng.random((K * (I + J), Q)).astype(T))
        self.U = np.asfortranarray(rng.random((K * (I + J), Q)).astype(T))
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
