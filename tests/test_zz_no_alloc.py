"""admx_loglikelihood_noalloc: the log-likelihood pass for callers that cannot
allocate (no reference counterpart).  Its inputs and output are caller-held
device buffers and its scratch is a fixed device array, so its result must
equal admx_loglikelihood_loop's on a context (same kernels, same chunking) bit
for bit, over consecutive calls that share that array: a degenerate state
(NaN log-likelihood, the state-check flag set) followed by a smaller cohort
in a normal state (the flag must start cleared again)."""
import ctypes as C

import numpy as np
import pytest

from conftest import make_cohort, random_state

H2D, D2H = 1, 2


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return hip


class DevBuf:
    """Device buffer owned by the test (the entry point itself allocates nothing)."""

    def __init__(self, hip, nbytes):
        self.hip, self.n = hip, nbytes
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), nbytes) == 0

    def put(self, a):
        a = np.ascontiguousarray(a) if a.ndim <= 1 else a
        assert a.nbytes <= self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, a):
        assert self.hip.hipMemcpy(a.ctypes.data, self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def _ll_class(v):
    return "nan" if np.isnan(v) else ("-inf" if v == -np.inf else "finite")


@pytest.mark.gpu
def test_noalloc_matches_context_entry_point(oa, O):
    from openadmixture_jl_amd import _native as N

    hip = _hip()
    cases = [(3000, 2500, 5, 1, True), (700, 300, 3, 2, False)]  # the second call is smaller
    for I, J, K, seed, degenerate in cases:
        print("noalloc", I, J, K, "degenerate" if degenerate else "normal", flush=True)
        Qt, Pt, bed, pitch0 = make_cohort(O, I, J, K, miss=0.02, seed=seed)
        q, p = random_state(I, J, K, seed)
        if degenerate:
            p[:, 7] = 0.0  # s = 0 on SNP 7: the reference's terms give NaN (0 log 0) or -Inf
            p = np.asfortranarray(p)
        c = oa.HIPAdmixData(I, J, K, 3)
        c.upload_bed(bed.reshape((pitch0, J), order="F"))
        want = c.loglikelihood_loop(q, p, 0, I, 0, J)
        c.close()
        # caller-held device buffers: BED rows at a 16-byte pitch, 32 readable bytes past the last row
        pitch = (pitch0 + 15) // 16 * 16
        rows = np.zeros((pitch, J), np.uint8, order="F")
        rows[:pitch0, :] = bed.reshape((pitch0, J), order="F")
        host_bed = np.concatenate([rows.ravel(order="F"), np.full(32, 0x55, np.uint8)])
        dbed, dq, dp, dout = (DevBuf(hip, host_bed.nbytes), DevBuf(hip, q.nbytes), DevBuf(hip, p.nbytes),
                              DevBuf(hip, 8))
        dbed.put(host_bed)
        dq.put(q.ravel(order="F"))
        dp.put(p.ravel(order="F"))
        rc = N.lib().admx_loglikelihood_noalloc(dbed.p, pitch, I, J, K, dq.p, dp.p, dout.p, None)
        assert rc == 0, N.lib().admx_last_error()
        assert hip.hipDeviceSynchronize() == 0
        got = float(dout.get(np.zeros(1))[0])
        for b in (dbed, dq, dp, dout):
            b.free()
        ref = O.loglikelihood(bed, pitch0, q, p)
        print(f"  noalloc {got!r} context {want!r} oracle {ref!r}", flush=True)
        assert _ll_class(got) == _ll_class(want) == _ll_class(ref)
        if degenerate:
            assert _ll_class(got) != "finite"
        else:
            assert got == want  # bit for bit: same kernels, same chunking, same reduction order
            assert abs(got - ref) <= 1e-10 * abs(ref)


def test_noalloc_refuses_cohorts_beyond_its_scratch(oa):
    """Argument checks run before any device work: a cohort whose partial sums
    would not fit the fixed scratch, a bad pitch and null pointers are refused
    with ADMX_EINVAL and a message."""
    from openadmixture_jl_amd import _native as N

    L = N.lib()
    fake = C.c_void_p(16)  # never dereferenced: the checks fail first
    too_many = 256 * (N.ADMX_NOALLOC_LL_PARTS + 1)  # ceil(I/256) partial blocks > the scratch, with one chunk
    assert L.admx_loglikelihood_noalloc(fake, (too_many + 3) // 4 + 16, too_many, 1, 2, fake, fake, fake,
                                        None) == N.ADMX_EINVAL
    assert b"ADMX_NOALLOC_LL_PARTS" in L.admx_last_error()
    assert L.admx_loglikelihood_noalloc(fake, 24, 64, 10, 2, fake, fake, fake, None) == N.ADMX_EINVAL  # pitch % 16
    assert L.admx_loglikelihood_noalloc(None, 16, 64, 10, 2, fake, fake, fake, None) == N.ADMX_EINVAL
