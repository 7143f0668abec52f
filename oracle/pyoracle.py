"""ctypes front-end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``; never by the product package.

Arrays follow the reference's Julia layout: column-major (``order='F'``)
K x I / K x J matrices, K x K x n Hessian stacks, SNP-major BED bytes with a
row pitch (bytes per SNP).  Ranges are 0-based half-open.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

c_i64 = C.c_int64
c_dbl = C.c_double
c_vp = C.c_void_p


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _declare(_lib)
    return _lib


def _sfx(dtype) -> str:
    return "_f64" if np.dtype(dtype) == np.float64 else "_f32"


def _declare(L):
    L.orc_decode.argtypes = [c_vp, c_i64, c_i64, c_i64, c_vp]
    L.orc_counts.argtypes = [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp]
    L.orc_gen_bed.argtypes = [c_vp, c_vp, c_i64, c_i64, c_i64, C.c_uint64, c_dbl, c_vp, c_i64]
    L.orc_fold_of.argtypes = [c_i64, c_i64, c_i64, C.c_uint64, C.c_int]
    L.orc_fold_of.restype = C.c_int
    L.orc_mask_fold.argtypes = [c_vp, c_vp, c_i64, c_i64, c_i64, C.c_int, C.c_int, C.c_uint64]
    L.orc_heldout_deviance.argtypes = [c_vp, c_i64, c_vp, c_vp, c_i64, c_i64, c_i64, C.c_int, C.c_int, C.c_uint64,
                                       c_vp]
    L.orc_heldout_deviance.restype = c_dbl
    for s in ("_f64", "_f32"):
        getattr(L, "orc_loglikelihood" + s).argtypes = [c_vp, c_i64, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_vp,
                                                        C.c_int]
        getattr(L, "orc_loglikelihood" + s).restype = c_dbl
        getattr(L, "orc_em_q_loop" + s).argtypes = [c_vp, c_vp, c_i64, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_vp,
                                                    C.c_int]
        getattr(L, "orc_em_p_loop" + s).argtypes = [c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_i64,
                                                    c_vp, C.c_int]
        for n in ("orc_update_q_loop", "orc_update_p_loop"):
            getattr(L, n + s).argtypes = [c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_vp,
                                          C.c_int]
        getattr(L, "orc_sweep" + s).argtypes = [c_vp, c_i64, c_i64, c_vp, C.c_int]
        fl = c_dbl if s == "_f64" else C.c_float
        getattr(L, "orc_create_tableau" + s).argtypes = [c_vp, c_vp, c_vp, c_vp, c_i64, C.c_int, fl]
        getattr(L, "orc_quadratic_program" + s).argtypes = [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_vp]
        getattr(L, "orc_quadratic_program" + s).restype = C.c_int
        getattr(L, "orc_project_q" + s).argtypes = [c_vp, c_i64, c_i64]
        getattr(L, "orc_project_p" + s).argtypes = [c_vp, c_i64]
        getattr(L, "orc_update_UV" + s).argtypes = [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64]
        getattr(L, "orc_update_QN" + s).argtypes = [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64]
        getattr(L, "orc_em_q" + s).argtypes = [c_vp]
        getattr(L, "orc_em_p" + s).argtypes = [c_vp]
        getattr(L, "orc_em_q_mode" + s).argtypes = [c_vp, C.c_int]
        getattr(L, "orc_em_p_mode" + s).argtypes = [c_vp, C.c_int]
        getattr(L, "orc_init_em" + s).argtypes = [c_vp, c_i64, c_vp]
        getattr(L, "orc_update_q" + s).argtypes = [c_vp, C.c_int, c_vp]
        getattr(L, "orc_update_p" + s).argtypes = [c_vp, C.c_int, c_vp]
        getattr(L, "orc_qn_iteration" + s).argtypes = [c_vp, c_i64, C.c_int, C.c_int, c_vp]
        getattr(L, "orc_qn_iteration_mode" + s).argtypes = [c_vp, c_i64, C.c_int, C.c_int, C.c_int, c_vp]
        getattr(L, "orc_update_UV_LBQN" + s).argtypes = [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64]
        getattr(L, "orc_update_QN_LBQN" + s).argtypes = [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64]
        for n in ("orc_qp_update_q", "orc_qp_update_p"):
            getattr(L, n + s).argtypes = [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_vp, C.c_int]


def _p(a):
    return None if a is None else a.ctypes.data_as(c_vp)


def _chk(a, dtype=None):
    assert a.flags.f_contiguous or a.ndim <= 1, "arrays must be column-major"
    if dtype is not None:
        assert a.dtype == np.dtype(dtype), (a.dtype, dtype)
    return a


def _nthreads(n):
    if n is not None:
        return int(n)
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


# ---------------------------------------------------------------- decode / counts
def decode(bed, pitch, I, J):
    g = np.empty((I, J), dtype=np.int8, order="F")
    lib().orc_decode(_p(bed), pitch, I, J, _p(g))
    return g


def counts(bed, pitch, I, J):
    snm = np.zeros(J, np.int64)
    salt = np.zeros(J, np.int64)
    inm = np.zeros(I, np.int64)
    lib().orc_counts(_p(bed), pitch, I, J, _p(snm), _p(salt), _p(inm))
    return snm, salt, inm


def gen_bed(Qt, Pt, seed, miss_rate, pitch=None):
    K, I = Qt.shape
    J = Pt.shape[1]
    Qt = np.asfortranarray(Qt, dtype=np.float64)
    Pt = np.asfortranarray(Pt, dtype=np.float64)
    if pitch is None:
        pitch = (I + 3) // 4
    bed = np.zeros(pitch * J, dtype=np.uint8)
    lib().orc_gen_bed(_p(Qt), _p(Pt), I, J, K, C.c_uint64(seed), miss_rate, _p(bed), pitch)
    return bed


# ---------------------------------------------------------------- cross-validation holdout
def fold_of(i, j, I, seed, nfolds):
    return lib().orc_fold_of(int(i), int(j), int(I), C.c_uint64(seed), int(nfolds))


def mask_fold(bed, pitch, I, J, fold, nfolds, seed):
    out = np.empty_like(bed)
    lib().orc_mask_fold(_p(bed), _p(out), pitch, I, J, int(fold), int(nfolds), C.c_uint64(seed))
    return out


def heldout_deviance(bed, pitch, q, p, fold, nfolds, seed):
    K, I = q.shape
    J = p.shape[1]
    cnt = C.c_int64()
    dev = lib().orc_heldout_deviance(_p(bed), pitch, _p(_chk(q, np.float64)), _p(_chk(p, np.float64)), I, J, K,
                                     int(fold), int(nfolds), C.c_uint64(seed), C.byref(cnt))
    return dev, cnt.value


# ---------------------------------------------------------------- tile loops
# mu: None for the reference's *_skipmissing loops; a length-J array of SNP
# means selects the mean-imputation loops (missing g = mu[j]).
def _mu(mu, dtype):
    return None if mu is None else np.ascontiguousarray(mu, dtype=dtype)


def snp_means(bed, pitch, I, J, dtype=np.float64):
    """SnpLinAlg's mean genotype per SNP over the observed entries (additive model)."""
    snm, salt, _ = counts(bed, pitch, I, J)
    return (salt / np.maximum(snm, 1)).astype(dtype)


def loglikelihood(bed, pitch, q, p, irange=None, jrange=None, nthreads=None, mu=None):
    K, I = q.shape
    J = p.shape[1]
    i0, i1 = irange if irange else (0, I)
    j0, j1 = jrange if jrange else (0, J)
    mu_a = _mu(mu, q.dtype)
    f = getattr(lib(), "orc_loglikelihood" + _sfx(q.dtype))
    return f(_p(bed), pitch, _p(_chk(q)), _p(_chk(p, q.dtype)), K, i0, i1, j0, j1, _p(mu_a), _nthreads(nthreads))


def em_q_loop(q_next, bed, pitch, q, p, irange=None, jrange=None, nthreads=None, mu=None):
    K, I = q.shape
    J = p.shape[1]
    i0, i1 = irange if irange else (0, I)
    j0, j1 = jrange if jrange else (0, J)
    mu_a = _mu(mu, q.dtype)
    getattr(lib(), "orc_em_q_loop" + _sfx(q.dtype))(_p(_chk(q_next, q.dtype)), _p(bed), pitch, _p(q), _p(p), K, i0, i1, j0, j1, _p(mu_a), _nthreads(nthreads))


def em_p_loop(p_next, p_tmp, bed, pitch, q, p, irange=None, jrange=None, nthreads=None, mu=None):
    K, I = q.shape
    J = p.shape[1]
    i0, i1 = irange if irange else (0, I)
    j0, j1 = jrange if jrange else (0, J)
    mu_a = _mu(mu, q.dtype)
    getattr(lib(), "orc_em_p_loop" + _sfx(q.dtype))(_p(_chk(p_next, q.dtype)), _p(_chk(p_tmp, q.dtype)), _p(bed), pitch, _p(q), _p(p), K, i0, i1, j0, j1, _p(mu_a), _nthreads(nthreads))


def update_q_loop(XtX, Xtz, bed, pitch, q, p, irange=None, jrange=None, nthreads=None, mu=None):
    K, I = q.shape
    J = p.shape[1]
    i0, i1 = irange if irange else (0, I)
    j0, j1 = jrange if jrange else (0, J)
    mu_a = _mu(mu, q.dtype)
    getattr(lib(), "orc_update_q_loop" + _sfx(q.dtype))(_p(_chk(XtX, q.dtype)), _p(_chk(Xtz, q.dtype)), _p(bed), pitch, _p(q), _p(p), K, i0, i1, j0, j1, _p(mu_a), _nthreads(nthreads))


def update_p_loop(XtX, Xtz, bed, pitch, q, p, irange=None, jrange=None, nthreads=None, mu=None):
    K, I = q.shape
    J = p.shape[1]
    i0, i1 = irange if irange else (0, I)
    j0, j1 = jrange if jrange else (0, J)
    mu_a = _mu(mu, q.dtype)
    getattr(lib(), "orc_update_p_loop" + _sfx(q.dtype))(_p(_chk(XtX, q.dtype)), _p(_chk(Xtz, q.dtype)), _p(bed), pitch, _p(q), _p(p), K, i0, i1, j0, j1, _p(mu_a), _nthreads(nthreads))


# ---------------------------------------------------------------- QP / projections / QN
def sweep(A, k, inverse=False):
    n = A.shape[0]
    tmp = np.empty(n, A.dtype)
    getattr(lib(), "orc_sweep" + _sfx(A.dtype))(_p(_chk(A)), n, k, _p(tmp), int(inverse))
    return A


def create_tableau(H, r, x, simplex, mu=0.0):
    K = H.shape[0]
    sz = K + 2 if simplex else K + 1
    tab = np.zeros((sz, sz), dtype=H.dtype, order="F")
    getattr(lib(), "orc_create_tableau" + _sfx(H.dtype))(_p(tab), _p(_chk(H)), _p(r), _p(x), K, int(simplex), mu)
    return tab


def quadratic_program(tab, par, pmin, pmax, c):
    """Returns (delta, status).  ``tab`` is modified in place."""
    p = par.shape[0]
    delta = np.zeros(p, tab.dtype)
    d = np.zeros(tab.shape[0], tab.dtype)
    tmp = np.zeros(tab.shape[0], tab.dtype)
    swept = np.zeros(p, np.uint8)
    f = getattr(lib(), "orc_quadratic_program" + _sfx(tab.dtype))
    r = f(_p(delta), _p(_chk(tab)), _p(par), _p(pmin), _p(pmax), p, c, _p(d), _p(tmp), _p(swept))
    return delta, r


def project_q(b):
    K, n = b.shape
    getattr(lib(), "orc_project_q" + _sfx(b.dtype))(_p(_chk(b)), K, n)
    return b


def project_p(b):
    getattr(lib(), "orc_project_p" + _sfx(b.dtype))(_p(_chk(b)), b.size)
    return b


def update_UV(U, V, x, x_next, x_next2, it, Q):
    N = x.size
    getattr(lib(), "orc_update_UV" + _sfx(x.dtype))(_p(U), _p(V), _p(x), _p(x_next), _p(x_next2), N, it, Q)


def update_QN(x_out, x_mapped, x, U, V, m):
    N = x.size
    getattr(lib(), "orc_update_QN" + _sfx(x.dtype))(_p(x_out), _p(x_mapped), _p(x), _p(U), _p(V), N, m)


def update_UV_LBQN(U, V, x, x_next, x_next2, it, Q):
    N = x.size
    getattr(lib(), "orc_update_UV_LBQN" + _sfx(x.dtype))(_p(U), _p(V), _p(x), _p(x_next), _p(x_next2), N, it, Q)


def update_QN_LBQN(x_out, x, U, V, m):
    N = x.size
    getattr(lib(), "orc_update_QN_LBQN" + _sfx(x.dtype))(_p(x_out), _p(x), _p(U), _p(V), N, m)


def qp_update(XtX, Xtz, x, x_next, simplex, nthreads=None):
    """The host half of update_q! (simplex=True) or update_p! (src/algorithms_inner.jl:197-230,
    273-307): negates Xtz in place, solves one QP per column of x (K x n) and writes
    the projected x + delta into x_next; returns the QP return codes."""
    K, n = x.shape
    diff = np.zeros((K, n), x.dtype, order="F")
    status = np.zeros(n, np.int32)
    f = getattr(lib(), ("orc_qp_update_q" if simplex else "orc_qp_update_p") + _sfx(x.dtype))
    f(_p(_chk(XtX, x.dtype)), _p(_chk(Xtz, x.dtype)), _p(_chk(x)), _p(_chk(x_next, x.dtype)), _p(diff), K, n,
      _p(status), _nthreads(nthreads))
    return status


# ---------------------------------------------------------------- whole algorithm
class _StateF64(C.Structure):
    _fields_ = [("I", c_i64), ("J", c_i64), ("K", c_i64), ("Q", c_i64),
                ("bed", c_vp), ("pitch", c_i64),
                ("x", c_vp), ("x_next", c_vp), ("x_next2", c_vp), ("x_tmp", c_vp),
                ("XtX_q", c_vp), ("Xtz_q", c_vp), ("XtX_p", c_vp), ("Xtz_p", c_vp),
                ("U", c_vp), ("V", c_vp),
                ("ll_prev", c_dbl), ("ll_new", c_dbl), ("nthreads", C.c_int), ("mu", c_vp)]


class OracleAdmix:
    """Host-array mirror of AdmixData{T} (src/structs.jl:9-215) driven by the oracle."""

    def __init__(self, bed, pitch, I, J, K, Q, x0, U0=None, V0=None, dtype=np.float64, nthreads=None, mu=None):
        self.dtype = np.dtype(dtype)
        self.I, self.J, self.K, self.Q = I, J, K, Q
        self.bed = np.ascontiguousarray(bed, dtype=np.uint8)
        self.pitch = pitch
        N = K * (I + J)
        self.x = np.asfortranarray(x0, dtype=self.dtype).copy(order="F")
        self.x_next = np.zeros((K, I + J), self.dtype, order="F")
        self.x_next2 = np.zeros((K, I + J), self.dtype, order="F")
        self.x_tmp = np.zeros((K, I + J), self.dtype, order="F")
        self.XtX_q = np.zeros((K, K, I), self.dtype, order="F")
        self.Xtz_q = np.zeros((K, I), self.dtype, order="F")
        self.XtX_p = np.zeros((K, K, J), self.dtype, order="F")
        self.Xtz_p = np.zeros((K, J), self.dtype, order="F")
        self.U = np.zeros((N, Q), self.dtype, order="F") if U0 is None else np.asfortranarray(U0, dtype=self.dtype).copy(order="F")
        self.V = np.zeros((N, Q), self.dtype, order="F") if V0 is None else np.asfortranarray(V0, dtype=self.dtype).copy(order="F")
        self.st = _StateF64(I, J, K, Q, _p(self.bed), pitch, _p(self.x), _p(self.x_next), _p(self.x_next2), _p(self.x_tmp),
                            _p(self.XtX_q), _p(self.Xtz_q), _p(self.XtX_p), _p(self.Xtz_p), _p(self.U), _p(self.V),
                            float("nan"), float("nan"), _nthreads(nthreads), None)
        if mu is not None:
            self.mu = np.ascontiguousarray(mu, dtype=self.dtype)
            self.st.mu = _p(self.mu)
        self.sfx = _sfx(self.dtype)

    @property
    def ll_new(self):
        return self.st.ll_new

    @ll_new.setter
    def ll_new(self, v):
        self.st.ll_new = v

    @property
    def q(self):
        return self.x[:, : self.I]

    @property
    def p(self):
        return self.x[:, self.I:]

    def _f(self, n):
        return getattr(lib(), n + self.sfx)

    def loglikelihood(self, x=None):
        x = self.x if x is None else x
        return loglikelihood(self.bed, self.pitch, np.asfortranarray(x[:, : self.I]), np.asfortranarray(x[:, self.I:]), nthreads=self.st.nthreads)

    EM_MODES = {"base": 0, "fallback": 1, "fallback2": 2}

    def em_q(self, mode="base"):
        self._f("orc_em_q_mode")(C.byref(self.st), self.EM_MODES[mode])

    def em_p(self, mode="base"):
        self._f("orc_em_p_mode")(C.byref(self.st), self.EM_MODES[mode])

    def init_em(self, iters):
        lls = np.zeros(max(iters, 1))
        self._f("orc_init_em")(C.byref(self.st), iters, _p(lls))
        return lls[:iters]

    def update_q(self, update2=False):
        status = np.zeros(self.I, np.int32)
        self._f("orc_update_q")(C.byref(self.st), int(update2), _p(status))
        return status

    def update_p(self, update2=False):
        status = np.zeros(self.J, np.int32)
        self._f("orc_update_p")(C.byref(self.st), int(update2), _p(status))
        return status

    def qn_iteration(self, it, fix_q=False, fix_p=False, mode="ZAL"):
        out = np.zeros(5)
        self._f("orc_qn_iteration_mode")(C.byref(self.st), it, int(fix_q), int(fix_p), 1 if mode == "LBQN" else 0,
                                         _p(out))
        return dict(prev=out[0], basic=out[1], qn=out[2], new=out[3], reldiff=out[4])


# ---------------------------------------------------------------- SKFR (oracle/skfr_oracle.c)
def _declare_skfr(L):
    L.orc_skfr_u.argtypes = [C.c_uint64, C.c_int]
    L.orc_skfr_u.restype = c_dbl
    L.orc_skfr_nsplit.argtypes = [c_i64, c_i64, C.c_int, C.c_int]
    L.orc_skfr_nsplit.restype = C.c_int
    L.orc_skfr_stats.argtypes = [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp]
    L.orc_skfr_init.argtypes = [c_vp, c_i64, c_i64, c_i64, C.c_int, C.c_uint64, C.c_int, c_vp, c_vp]
    L.orc_skfr_init.restype = C.c_int
    L.orc_skfr_run.argtypes = [c_vp, c_i64, c_i64, c_i64, C.c_int, C.c_int, c_i64, C.c_int, C.c_int, c_vp, c_vp, c_vp,
                               c_vp, c_vp, c_vp, c_vp]
    L.orc_skfr_run.restype = C.c_int


def _skfr_lib():
    L = lib()
    if not hasattr(L, "_skfr_declared"):
        _declare_skfr(L)
        L._skfr_declared = True
    return L


def _bedp(bed, I, J):
    """(pitch x J Fortran BED, pitch) from a 2-D payload or a flat one of J equal rows."""
    bed = np.asarray(bed, dtype=np.uint8)
    if bed.ndim == 1:
        if J < 1 or bed.size % J:
            raise ValueError("flat BED size is not a multiple of J")
        bed = bed.reshape((bed.size // J, J), order="F")
    bed = np.asfortranarray(bed)
    if bed.shape[1] != J or bed.shape[0] < (I + 3) // 4:
        raise ValueError(f"BED of shape {bed.shape} does not hold I={I}, J={J}")
    return bed, bed.shape[0]


def skfr_u(seed: int, cc: int) -> float:
    return _skfr_lib().orc_skfr_u(C.c_uint64(seed), cc)


def skfr_nsplit(I: int, ncols: int, nsplit: int, R: int = 1) -> int:
    return _skfr_lib().orc_skfr_nsplit(I, ncols, nsplit, R)


def skfr_stats(bed, I: int, J: int):
    """(mu_j, v_j) of a pitch x J BED: mean over observed entries, mu (1 - mu/2)."""
    bed, pitch = _bedp(bed, I, J)
    mu, var = np.zeros(J), np.zeros(J)
    _skfr_lib().orc_skfr_stats(bed.ctypes.data, pitch, I, J, mu.ctypes.data, var.ctypes.data)
    return mu, var


def skfr_init(bed, I: int, J: int, k: int, seed: int, nsplit: int):
    """k-means++ start: (clusters, drawn center individuals)."""
    bed, pitch = _bedp(bed, I, J)
    cl = np.zeros(I, np.int32)
    cen = np.zeros(k, np.int64)
    rc = _skfr_lib().orc_skfr_init(bed.ctypes.data, pitch, I, J, k, C.c_uint64(seed), nsplit, cl.ctypes.data,
                                   cen.ctypes.data)
    if rc:
        raise ValueError("orc_skfr_init: bad arguments")
    return cl, cen


def skfr_run(bed, I: int, J: int, k: int, clusters, sparsity: int, mode: str = "global", max_iter: int = 50,
             nsplit: int = 1):
    """Sparse k-means by feature ranking: (clusters, centers k x J, aims, WSS, obj, iters)."""
    bed, pitch = _bedp(bed, I, J)
    cl = np.ascontiguousarray(clusters, dtype=np.int32).copy()
    if cl.shape != (I,) or cl.min() < 0 or cl.max() >= k:
        raise ValueError("clusters must hold I labels in [0, k)")
    local = 1 if mode == "local" else 0
    cap = min(J, k * sparsity) if local else sparsity
    aims = np.zeros(max(cap, 1), np.int64)
    n_aims = C.c_int64()
    cen = np.zeros((k, J), order="F")
    wss = np.zeros(k)
    obj, iters = C.c_double(), C.c_int32()
    rc = _skfr_lib().orc_skfr_run(bed.ctypes.data, pitch, I, J, k, local, sparsity, max_iter, nsplit, cl.ctypes.data,
                                  aims.ctypes.data, C.addressof(n_aims), cen.ctypes.data, wss.ctypes.data,
                                  C.addressof(obj), C.addressof(iters))
    if rc:
        raise ValueError("orc_skfr_run: bad arguments")
    return cl, cen, aims[: n_aims.value].copy(), wss, obj.value, iters.value
