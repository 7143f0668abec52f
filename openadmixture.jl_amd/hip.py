"""Device context: the MI355X counterpart of ``CuAdmixData{T}`` and ``g_cu``
(src/cuda/structs.jl:1-38).

A ``HIPAdmixData`` owns, in HBM, the 2-bit BED payload, the four state vectors
x, x_next, x_next2, x_tmp (K x (I+J) each, q = first I columns), the Hessian
and gradient stacks of both sides, the quasi-Newton history U, V and the
log-likelihood slots.  Every method is one call into libadmixhip.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


def _f(a, dtype, shape=None):
    a = np.asarray(a)
    if a.dtype != dtype or not (a.flags.f_contiguous or a.ndim <= 1):
        raise TypeError(f"expected a column-major {np.dtype(dtype)} array, got {a.dtype} "
                        f"(f_contiguous={a.flags.f_contiguous})")
    if shape is not None and a.shape != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {a.shape}")
    return a


class HIPAdmixData:
    """Device mirror of AdmixData{T} plus the device-resident BED."""

    def __init__(self, I: int, J: int, K: int, Q: int = 3, T=np.float64, stream_snps: int = 0,
                 bed_file: str | None = None, resident_snps: int = 0):
        """stream_snps > 0: the genotypes stay in pinned host memory and every
        pass streams them through HBM in chunks of that many SNPs (cohorts
        larger than HBM).  With bed_file as well, they are read from that
        PLINK .bed file chunk by chunk instead (no host copy; cohorts larger
        than host memory), and the context needs no upload.  resident_snps:
        the first that many SNPs of a streamed context stay in HBM."""
        self.I, self.J, self.K, self.Q = int(I), int(J), int(K), int(Q)
        self.T = np.dtype(T)
        self.stream_snps = int(stream_snps)
        self.bed_file = bed_file
        self._h = C.c_void_p()
        if bed_file is not None:
            if not self.stream_snps:
                raise ValueError("a file-backed context streams: give stream_snps > 0")
            N.check(N.lib().admx_create_streamed_file(C.byref(self._h), str(bed_file).encode(), self.I, self.J,
                                                      self.K, self.Q, N.dtype_code(self.T), self.stream_snps))
        elif self.stream_snps:
            N.check(N.lib().admx_create_streamed(C.byref(self._h), self.I, self.J, self.K, self.Q,
                                                 N.dtype_code(self.T), self.stream_snps))
        else:
            N.check(N.lib().admx_create(C.byref(self._h), self.I, self.J, self.K, self.Q, N.dtype_code(self.T)))
        if resident_snps:
            self.set_resident_snps(resident_snps)

    # ------------------------------------------------------------ lifetime
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().admx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    # ------------------------------------------------------------ genotypes
    def upload_bed(self, data):
        """Raw SNP-major payload: a (bytes_per_snp x J) column-major uint8 array
        (SnpArray.data), or flat bytes with bytes_per_snp = ceil(I/4)."""
        data = np.asarray(data, dtype=np.uint8)
        if data.ndim == 2:
            if data.shape[1] != self.J:
                raise ValueError("BED payload must have J columns")
            bps = data.shape[0]
            data = np.asfortranarray(data)
        else:
            bps = (self.I + 3) // 4
            if data.size != bps * self.J:
                raise ValueError("flat BED payload must hold ceil(I/4)*J bytes")
        N.check(N.lib().admx_upload_bed(self._h, N.ptr(data), bps))

    def download_bed(self):
        out = np.empty(((self.I + 3) // 4, self.J), dtype=np.uint8, order="F")
        N.check(N.lib().admx_download_bed(self._h, N.ptr(out)))
        return out

    def generate_bed(self, Qt, Pt, seed: int, miss_rate: float):
        Qt = np.asfortranarray(Qt, dtype=np.float64)
        Pt = np.asfortranarray(Pt, dtype=np.float64)
        if Qt.shape != (self.K, self.I) or Pt.shape != (self.K, self.J):
            raise ValueError("Q must be K x I and P must be K x J")
        N.check(N.lib().admx_generate_bed(self._h, N.ptr(Qt), N.ptr(Pt), C.c_uint64(seed), float(miss_rate)))

    def decode(self):
        g = np.empty((self.I, self.J), dtype=np.int8, order="F")
        N.check(N.lib().admx_decode(self._h, N.ptr(g)))
        return g

    def counts(self):
        snm = np.zeros(self.J, np.int64)
        salt = np.zeros(self.J, np.int64)
        inm = np.zeros(self.I, np.int64)
        N.check(N.lib().admx_counts(self._h, N.ptr(snm), N.ptr(salt), N.ptr(inm)))
        return snm, salt, inm

    # ------------------------------------------------------------ cross-validation holdout
    def bed_from(self, src: "HIPAdmixData", fold: int = -1, nfolds: int = 5, seed: int = 0):
        """Genotypes of `src` (same I, J), with the entries of `fold` set to
        missing (fold < 0: plain copy); device to device."""
        N.check(N.lib().admx_bed_from(self._h, src.handle, int(fold), int(nfolds), C.c_uint64(seed)))

    def heldout_deviance(self, src: "HIPAdmixData", fold: int, nfolds: int = 5, seed: int = 0):
        """(deviance sum, entry count) of this context's state x on the observed
        entries of `fold` of src's genotypes."""
        dev, cnt = C.c_double(), C.c_int64()
        N.check(N.lib().admx_heldout_deviance(self._h, src.handle, int(fold), int(nfolds), C.c_uint64(seed),
                                              C.byref(dev), C.byref(cnt)))
        return dev.value, cnt.value

    # ------------------------------------------------------------ SKFR AIM pre-selection
    def skfr_init(self, k: int, seed: int = 0, nsplit: int = 0) -> np.ndarray:
        """k-means++ initial clusters (I, 0-based) of the standardized genotypes."""
        cl = np.zeros(self.I, np.int32)
        N.check(N.lib().admx_skfr_init(self._h, int(k), C.c_uint64(seed), int(nsplit), N.ptr(cl)))
        return cl

    def skfr_run(self, k: int, clusters, sparsity: int, mode: str = "global", max_iter: int = 50,
                 nsplit: int = 0, centers: bool = True):
        """Sparse k-means by feature ranking from `clusters`; returns
        (clusters, centers (k x J or None), aims (0-based, decreasing score),
        WSS (k), objective, iterations)."""
        modes = {"global": N.ADMX_SKFR_GLOBAL, "local": N.ADMX_SKFR_LOCAL}
        if mode not in modes:
            raise ValueError("skfr mode can only be 'global' or 'local'")
        cl = np.ascontiguousarray(clusters, dtype=np.int32).copy()
        if cl.shape != (self.I,):
            raise ValueError(f"clusters must have {self.I} entries")
        cap = min(self.J, int(k) * int(sparsity)) if mode == "local" else int(sparsity)
        aims = np.zeros(max(cap, 1), np.int64)
        n_aims = C.c_int64()
        cen = np.zeros((int(k), self.J), np.float64, order="F") if centers else None
        wss = np.zeros(int(k), np.float64)
        obj, iters = C.c_double(), C.c_int32()
        N.check(N.lib().admx_skfr_run(self._h, int(k), modes[mode], int(sparsity), int(max_iter), int(nsplit),
                                      N.ptr(cl), N.ptr(aims), C.byref(n_aims), N.ptr(cen), N.ptr(wss),
                                      C.byref(obj), C.byref(iters)))
        return cl, cen, aims[: n_aims.value].copy(), wss, obj.value, iters.value

    # ------------------------------------------------------------ learned warm start
    def features(self, R, mu, isd, out=None):
        """Genotype-projection features over the first len(mu) SNPs (kept on the
        device for warm_q); returns them when `out` is given."""
        Jp = int(np.asarray(mu).size)
        N.check(N.lib().admx_features(self._h, N.ptr(R), N.ptr(mu), N.ptr(isd), Jp, N.ptr(out)))
        return out

    def warm_q(self, net):
        """q of the state x <- projected network output on the stored features."""
        N.check(N.lib().admx_warm_q(self._h, net.handle))

    def warm_p(self) -> float:
        """p of the state x <- observed allele frequencies (clipped), on the device.
        Returns the log-likelihood at that p, which every q on the simplex gives
        (Hardy-Weinberg: s = q.p = f), summed from the same per-SNP counts."""
        out = C.c_double()
        N.check(N.lib().admx_warm_p(self._h, C.byref(out)))
        return out.value

    def copy_x(self, dst: int, src: int):
        """State vector dst <- state vector src, on the device."""
        N.check(N.lib().admx_copy_x(self._h, int(dst), int(src)))

    # ------------------------------------------------------------ state
    def set_x(self, which: int, x):
        x = _f(x, self.T, (self.K, self.I + self.J))
        N.check(N.lib().admx_set_x(self._h, which, N.ptr(x)))

    def get_x(self, which: int, out=None):
        if out is None:
            out = np.empty((self.K, self.I + self.J), self.T, order="F")
        _f(out, self.T, (self.K, self.I + self.J))
        N.check(N.lib().admx_get_x(self._h, which, N.ptr(out)))
        return out

    # ------------------------------------------------------------ per-array mirror of CuAdmixData
    def _block_shape(self, side):
        return (self.K, self.I if side == N.ADMX_SIDE_Q else self.J)

    def set_block(self, which: int, side: int, a):
        """copyto!(d_cu.q / d_cu.p, a): the q (side Q) or p (side P) block of
        state vector `which` from a host array."""
        a = _f(a, self.T, self._block_shape(side))
        N.check(N.lib().admx_set_block(self._h, int(which), int(side), N.ptr(a)))

    def get_block(self, which: int, side: int, out=None):
        """copyto!(host, d_cu.q / d_cu.p)."""
        if out is None:
            out = np.empty(self._block_shape(side), self.T, order="F")
        _f(out, self.T, self._block_shape(side))
        N.check(N.lib().admx_get_block(self._h, int(which), int(side), N.ptr(out)))
        return out

    def copy_block(self, dst: int, src: int, side: int):
        """d_cu.p .= d_cu.p_next (side P) / d_cu.q .= d_cu.q_next (side Q), on the device."""
        N.check(N.lib().admx_copy_block(self._h, int(dst), int(src), int(side)))

    def grad_hessian(self, side: int):
        """update_q_cuda! / update_p_cuda!: XtX, Xtz of state x's (q, p) on `side`
        into the context (read them with hessians(side); Xtz not negated)."""
        N.check(N.lib().admx_grad_hessian(self._h, int(side)))

    def set_uv(self, U, V):
        shp = (self.K * (self.I + self.J), self.Q)
        N.check(N.lib().admx_set_uv(self._h, N.ptr(_f(U, self.T, shp)), N.ptr(_f(V, self.T, shp))))

    def get_uv(self, U=None, V=None):
        shp = (self.K * (self.I + self.J), self.Q)
        U = np.empty(shp, self.T, order="F") if U is None else _f(U, self.T, shp)
        V = np.empty(shp, self.T, order="F") if V is None else _f(V, self.T, shp)
        N.check(N.lib().admx_get_uv(self._h, N.ptr(U), N.ptr(V)))
        return U, V

    def set_ll(self, ll_prev: float, ll_new: float):
        N.check(N.lib().admx_set_ll(self._h, float(ll_prev), float(ll_new)))

    def get_ll(self):
        a, b = C.c_double(), C.c_double()
        N.check(N.lib().admx_get_ll(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def hessians(self, side: int):
        n = self.I if side == N.ADMX_SIDE_Q else self.J
        XtX = np.empty((self.K, self.K, n), self.T, order="F")
        Xtz = np.empty((self.K, n), self.T, order="F")
        N.check(N.lib().admx_get_hessians(self._h, side, N.ptr(XtX), N.ptr(Xtz)))
        return XtX, Xtz

    # ------------------------------------------------------------ device-resident algorithms
    def loglikelihood(self, which: int = N.ADMX_X) -> float:
        out = C.c_double()
        N.check(N.lib().admx_loglikelihood(self._h, which, C.byref(out)))
        return out.value

    def loglikelihood2(self, which1: int, which2: int):
        """Log-likelihoods of two state vectors, from one pass over the
        genotypes where admixture_qn! pairs them too (each identical to
        loglikelihood's)."""
        out = (C.c_double * 2)()
        N.check(N.lib().admx_loglikelihood2(self._h, which1, which2, out))
        return out[0], out[1]

    EM_MODES = {"base": N.ADMX_EM_BASE, "fallback": N.ADMX_EM_FALLBACK, "fallback2": N.ADMX_EM_FALLBACK2}

    def em_q(self, mode: str = "base"):
        """em_q!(d, g; mode): base/fallback q_next from (q, p); fallback2 q_next2
        from (q_next, p_next)."""
        if mode == "base":
            N.check(N.lib().admx_em_q(self._h))
        else:
            N.check(N.lib().admx_em_q_mode(self._h, self._em_mode(mode)))

    def em_p(self, mode: str = "base"):
        """em_p!(d, g; mode): base p_next from (q, p); fallback p_next from
        (q_next, p); fallback2 p_next2 from (q_next2, p_next)."""
        if mode == "base":
            N.check(N.lib().admx_em_p(self._h))
        else:
            N.check(N.lib().admx_em_p_mode(self._h, self._em_mode(mode)))

    def _em_mode(self, mode):
        m = str(mode).lstrip(":")
        if m not in self.EM_MODES:
            raise ValueError(f"em mode must be one of {sorted(self.EM_MODES)}")
        return self.EM_MODES[m]

    def init_em(self, iters: int):
        lls = np.zeros(max(int(iters), 1))
        N.check(N.lib().admx_init_em(self._h, int(iters), N.ptr(lls)))
        return lls[: int(iters)]

    def update_q(self, update2=False, status=False):
        st = np.zeros(self.I, np.int32) if status else None
        N.check(N.lib().admx_update_q(self._h, int(bool(update2)), N.ptr(st)))
        return st

    def update_p(self, update2=False, status=False):
        st = np.zeros(self.J, np.int32) if status else None
        N.check(N.lib().admx_update_p(self._h, int(bool(update2)), N.ptr(st)))
        return st

    def qn_iteration(self, it: int, fix_q=False, fix_p=False):
        out = np.zeros(5)
        N.check(N.lib().admx_qn_iteration(self._h, int(it), int(bool(fix_q)), int(bool(fix_p)), N.ptr(out)))
        return dict(prev=out[0], basic=out[1], qn=out[2], new=out[3], reldiff=out[4])

    def synchronize(self):
        N.check(N.lib().admx_synchronize(self._h))

    # ------------------------------------------------------------ host-array tile loops
    def _qp(self, q, p):
        return _f(q, self.T, (self.K, self.I)), _f(p, self.T, (self.K, self.J))

    def loglikelihood_loop(self, q, p, i0, i1, j0, j1) -> float:
        q, p = self._qp(q, p)
        out = C.c_double()
        N.check(N.lib().admx_loglikelihood_loop(self._h, N.ptr(q), N.ptr(p), i0, i1, j0, j1, C.byref(out)))
        return out.value

    def em_q_loop(self, q_next, q, p, i0, i1, j0, j1):
        q, p = self._qp(q, p)
        _f(q_next, self.T, (self.K, self.I))
        N.check(N.lib().admx_em_q_loop(self._h, N.ptr(q_next), N.ptr(q), N.ptr(p), i0, i1, j0, j1))

    def em_p_loop(self, p_next, p_tmp, q, p, i0, i1, j0, j1):
        q, p = self._qp(q, p)
        _f(p_next, self.T, (self.K, self.J))
        _f(p_tmp, self.T, (self.K, self.J))
        N.check(N.lib().admx_em_p_loop(self._h, N.ptr(p_next), N.ptr(p_tmp), N.ptr(q), N.ptr(p), i0, i1, j0, j1))

    def update_q_loop(self, XtX, Xtz, q, p, i0, i1, j0, j1):
        q, p = self._qp(q, p)
        _f(XtX, self.T, (self.K, self.K, self.I))
        _f(Xtz, self.T, (self.K, self.I))
        N.check(N.lib().admx_update_q_loop(self._h, N.ptr(XtX), N.ptr(Xtz), N.ptr(q), N.ptr(p), i0, i1, j0, j1))

    def update_p_loop(self, XtX, Xtz, q, p, i0, i1, j0, j1):
        q, p = self._qp(q, p)
        _f(XtX, self.T, (self.K, self.K, self.J))
        _f(Xtz, self.T, (self.K, self.J))
        N.check(N.lib().admx_update_p_loop(self._h, N.ptr(XtX), N.ptr(Xtz), N.ptr(q), N.ptr(p), i0, i1, j0, j1))

    def qp_batch(self, simplex: bool, XtX, Xtz_neg, x):
        n = x.shape[1]
        _f(XtX, self.T, (self.K, self.K, n))
        _f(Xtz_neg, self.T, (self.K, n))
        _f(x, self.T, (self.K, n))
        delta = np.zeros((self.K, n), self.T, order="F")
        status = np.zeros(n, np.int32)
        N.check(N.lib().admx_qp_batch(self._h, int(bool(simplex)), n, N.ptr(XtX), N.ptr(Xtz_neg), N.ptr(x),
                                      N.ptr(delta), N.ptr(status)))
        return delta, status

    def project_q(self, b):
        _f(b, self.T)
        N.check(N.lib().admx_project_q(self._h, N.ptr(b), b.size // self.K))
        return b

    def project_p(self, b):
        _f(b, self.T)
        N.check(N.lib().admx_project_p(self._h, N.ptr(b), b.size))
        return b

    def zal_step(self, it, x, x_next, x_next2, U, V, x_tmp):
        shp = (self.K, self.I + self.J)
        args = [_f(a, self.T, shp) for a in (x, x_next, x_next2)]
        ushp = (self.K * (self.I + self.J), self.Q)
        _f(U, self.T, ushp)
        _f(V, self.T, ushp)
        _f(x_tmp, self.T, shp)
        N.check(N.lib().admx_zal_step(self._h, int(it), *[N.ptr(a) for a in args], N.ptr(U), N.ptr(V),
                                      N.ptr(x_tmp)))

    def lbqn_step(self, it, x, x_next, x_next2, U, V, x_tmp):
        shp = (self.K, self.I + self.J)
        args = [_f(a, self.T, shp) for a in (x, x_next, x_next2)]
        ushp = (self.K * (self.I + self.J), self.Q)
        _f(U, self.T, ushp)
        _f(V, self.T, ushp)
        _f(x_tmp, self.T, shp)
        N.check(N.lib().admx_lbqn_step(self._h, int(it), *[N.ptr(a) for a in args], N.ptr(U), N.ptr(V),
                                       N.ptr(x_tmp)))

    def set_skipmissing(self, skip: bool = True):
        """True: the *_skipmissing loops; False: the mean-imputation loops (a
        missing genotype counts as its SNP's mean over the observed entries)."""
        self.set_tuning(N.ADMX_TUNE_SKIPMISSING, int(bool(skip)))
        self._skipmissing = bool(skip)

    def set_qn_mode(self, mode: str):
        """admixture_qn!'s quasi-Newton step: 'ZAL' or 'LBQN'."""
        codes = {"ZAL": N.ADMX_QN_ZAL, "LBQN": N.ADMX_QN_LBQN}
        if mode not in codes:
            raise ValueError("mode must be 'ZAL' or 'LBQN'")
        self.set_tuning(N.ADMX_TUNE_QN_MODE, codes[mode])

    # ------------------------------------------------------------ measurement / tuning
    def profile(self, on: bool = True):
        N.check(N.lib().admx_profile_enable(self._h, int(bool(on))))

    def profile_reset(self):
        N.check(N.lib().admx_profile_reset(self._h))

    def profile_report(self) -> dict:
        buf = C.create_string_buffer(1 << 16)
        N.check(N.lib().admx_profile_report(self._h, buf, len(buf)))
        rep = {}
        for line in buf.value.decode().splitlines():
            name, n, ms = line.split()
            rep[name] = (int(n), float(ms))
        return rep

    def set_precision(self, mode):
        """'f64' (plain), 'f32' or 'bf16' for the device-resident update passes."""
        codes = {"f64": N.ADMX_PREC_F64, "f32": N.ADMX_PREC_F32, "bf16": N.ADMX_PREC_BF16}
        self.set_tuning(N.ADMX_TUNE_PRECISION, codes[mode] if isinstance(mode, str) else int(mode))
        self._precision = mode

    def set_resident_snps(self, n: int):
        """Streamed context: keep SNPs [0, n) in HBM; passes stream only the
        rest.  n = -1: as many as the free HBM holds (a reserve stays free for
        the passes' scratch)."""
        self.set_tuning(N.ADMX_TUNE_RESIDENT_SNPS, int(n))

    def set_graphs(self, on: bool = True):
        """Replay steady-state qn_iteration calls from captured hipGraphs."""
        self.set_tuning(N.ADMX_TUNE_GRAPHS, int(bool(on)))

    def set_tuning(self, key: int, value: int):
        N.check(N.lib().admx_set_tuning(self._h, key, int(value)))
