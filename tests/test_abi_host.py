"""CPU-side checks of the boundary and the host mirror: the C ABI library loads
and exports every function include/admixhip.h declares, the Python binding
covers all of them, and the host-side logic (ranges, AdmixData, PLINK I/O,
argument validation) behaves as the reference's does.  No device calls."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "admixhip.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(admx_\w+)\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol(oa):
    names = header_functions()
    assert len(names) >= 30
    lib = oa._native.lib()
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(oa._native.SIGNATURES), set(names) ^ set(oa._native.SIGNATURES)


def test_nm_exports_match_header():
    so = os.path.join(ROOT, "openadmixture.jl_amd", "libadmixhip.so")
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (admx_\w+)$", out, flags=re.M))
    assert exported == set(header_functions())


def test_version_and_error_path(oa):
    lib = oa._native.lib()
    assert b"gfx950" in lib.admx_version()
    try:
        n = oa.device_count()
    except oa.AdmixHIPError as e:  # no device in this container: a clean error, no crash
        assert "libadmixhip error" in str(e)
        with pytest.raises(oa.AdmixHIPError):
            oa.HIPAdmixData(10, 10, 3)
    else:
        assert n >= 1


def test_create_rejects_bad_shapes(oa):
    import ctypes as C

    lib = oa._native.lib()
    h = C.c_void_p()
    assert lib.admx_create(C.byref(h), 10, 10, 17, 3, 0) == oa._native.ADMX_EINVAL
    assert b"K must be" in lib.admx_last_error()
    assert lib.admx_create(C.byref(h), 10, 10, 3, 9, 0) == oa._native.ADMX_EINVAL
    assert lib.admx_create(C.byref(h), 0, 10, 3, 3, 0) == oa._native.ADMX_EINVAL
    assert lib.admx_create(C.byref(h), 10, 10, 3, 3, 7) == oa._native.ADMX_EINVAL
    assert lib.admx_destroy(None) == 0


def test_missing_library_fails_loudly(oa, monkeypatch):
    N = oa._native
    monkeypatch.setattr(N, "_lib", None)
    monkeypatch.setattr(N, "LIB_PATH", "/nonexistent/libadmixhip.so")
    with pytest.raises(N.AdmixHIPError, match="not found"):
        N.lib()


def test_ranges(oa):
    from openadmixture_jl_amd.loops import _bounds, _ctx

    assert _bounds(oa.UnitRange(1, 10)) == (0, 10)
    assert _bounds(oa.UnitRange(3, 2)) == (2, 2)
    assert _bounds(range(4, 9)) == (4, 9)
    assert len(oa.UnitRange(5, 7)) == 3
    with pytest.raises(ValueError):
        _bounds(range(0, 10, 2))
    with pytest.raises(TypeError):
        _ctx(object())


def test_admixdata_layout(oa):
    d = oa.AdmixData(7, 11, 3, 3, rng=np.random.default_rng(0))
    assert d.x.shape == (3, 18) and d.x.flags.f_contiguous
    assert np.allclose(d.q.sum(0), 1)
    assert np.shares_memory(d.q, d.x) and np.shares_memory(d.p, d.x)
    d.p[0, 0] = 0.25
    assert d.x[0, 7] == 0.25
    assert d.U.shape == (3 * 18, 3) and d.XtX_q.shape == (3, 3, 7)
    assert np.isnan(d.ll_new)
    d32 = oa.AdmixData(5, 5, 2, 3, T=np.float32)
    assert d32.x.dtype == np.float32


def test_cpu_path_is_refused(oa):
    d = oa.AdmixData(5, 5, 2, 3, rng=np.random.default_rng(0))
    with pytest.raises(TypeError, match="no CPU path"):
        oa.update_q(d, None)
    with pytest.raises(TypeError):
        oa.admixture_qn(d, None, 1, log=None)


def test_plink_round_trip(oa, tmp_path):
    I, J = 13, 9
    rng = np.random.default_rng(1)
    data = rng.integers(0, 256, size=((I + 3) // 4, J), dtype=np.uint8)
    bed = oa.write_plink(str(tmp_path / "toy"), data, I, J)
    s = oa.SnpArray(bed)
    assert s.shape == (I, J)
    assert np.array_equal(np.asarray(s.data), data)
    g = oa.SnpLinAlg(s)
    assert g.shape == (I, J)
    with open(bed, "r+b") as f:
        f.write(b"\x00")
    with pytest.raises(ValueError, match="magic"):
        oa.SnpArray(bed)
    with pytest.raises(ValueError):
        oa.SnpArray(str(tmp_path / "toy.txt"))


def test_streamed_file_context_validates_the_bed(oa, tmp_path):
    """admx_create_streamed_file checks the file before touching the device:
    missing file -> ADMX_EIO; wrong magic or a size other than 3 + ceil(I/4)*J
    -> ADMX_EINVAL; stream_snps is required."""
    I, J = 13, 9
    data = np.random.default_rng(2).integers(0, 256, size=((I + 3) // 4, J), dtype=np.uint8)
    bed = oa.write_plink(str(tmp_path / "toy"), data, I, J)
    N = oa._native
    with pytest.raises(oa.AdmixHIPError, match=f"error {N.ADMX_EIO}: cannot open"):
        oa.HIPAdmixData(I, J, 3, 3, stream_snps=16, bed_file=str(tmp_path / "none.bed"))
    with pytest.raises(oa.AdmixHIPError, match=f"error {N.ADMX_EINVAL}: .*3 \\+ ceil"):
        oa.HIPAdmixData(I + 4, J, 3, 3, stream_snps=16, bed_file=bed)
    with pytest.raises(oa.AdmixHIPError, match=f"error {N.ADMX_EINVAL}: .*at least 16"):
        oa.HIPAdmixData(I, J, 3, 3, stream_snps=8, bed_file=bed)
    with pytest.raises(ValueError, match="stream_snps"):
        oa.HIPAdmixData(I, J, 3, 3, bed_file=bed)
    with open(bed, "r+b") as f:
        f.write(b"\x6c\x1b\x00")  # individual-major mode
    with pytest.raises(oa.AdmixHIPError, match=f"error {N.ADMX_EINVAL}: .*SNP-major"):
        oa.HIPAdmixData(I, J, 3, 3, stream_snps=16, bed_file=bed)


def test_run_admixture_argument_checks(oa, tmp_path):
    with pytest.raises(ValueError):
        oa.run_admixture(str(tmp_path / "x.txt"), 3, log=None)
    with pytest.raises(ValueError):
        oa.run_admixture(str(tmp_path / "x.bed"), 3, sparsity=100, skfr_mode="other", log=None)


def test_sim_truth(oa):
    from openadmixture_jl_amd import sim

    Qt, Pt = sim.draw_truth(50, 40, 4, 0.3, "beta", 0.01, 0.99, seed=2)
    assert Qt.shape == (4, 50) and Pt.shape == (4, 40)
    assert np.allclose(Qt.sum(0), 1)
    assert Pt.min() >= 0.01 and Pt.max() <= 0.99


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
