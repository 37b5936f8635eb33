"""The allocation-free entry point rmx_scale_l2_noalloc (through the C ABI) against the
existing rmx_scale_l2 on the same systems: the same kernels, so the scale, the scaled matrix
and the scaled right-hand side are bit-identical.  Two calls in a row, the second on a smaller
system (the static scratch keeps the first call's contents), in FP64 and FP32.  The test's
own device arrays come from the HIP runtime librmx.so links (opened without loading a
second copy), not from the entry point."""
import ctypes as C
import os

import numpy as np
import pytest

from rismo2d_amd import rmx

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    rmx.lib()                                    # loads librmx.so and the HIP runtime it links
    return C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD)


class _Dev:
    """A device array of the test (hipMalloc / hipMemcpy / hipFree)."""

    def __init__(self, hip, a):
        self.hip, self.a = hip, np.ascontiguousarray(a)
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), C.c_size_t(max(self.a.nbytes, 8))) == 0
        assert hip.hipMemcpy(self.p, self.a.ctypes.data_as(C.c_void_p), C.c_size_t(self.a.nbytes),
                             H2D) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(out.nbytes),
                                  D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _system(n, seed):
    """CSR with the diagonal first, random off-diagonal pattern and values."""
    rng = np.random.default_rng(seed)
    rp, cols, vals = [0], [], []
    for i in range(n):
        js = [i] + sorted(set(rng.integers(0, n, 6).tolist()) - {i})
        cols += js
        vals += (rng.standard_normal(len(js)) * 10.0).tolist()
        rp.append(len(cols))
    return (np.asarray(rp, np.int32), np.asarray(cols, np.int32), np.asarray(vals),
            rng.standard_normal(n) * 3.0)


def _noalloc(lib, n, d_rp, prec, d_A, d_B, d_s):
    f = lib.rmx_scale_l2_noalloc
    f.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_void_p]
    f.restype = C.c_int
    return f(0, n, d_rp, prec, d_A, d_B, d_s, None)


def test_noalloc_scale_matches_context_scale():
    print("[gpu] rmx_scale_l2_noalloc vs rmx_scale_l2, FP64 then FP32, two sizes", flush=True)
    hip = _hip()
    lib = rmx.lib()
    ctx = rmx.Context(0)
    try:
        for n, seed in ((5000, 1), (700, 2)):              # the second call is smaller
            rp, cols, vals, B = _system(n, seed)
            ctx.csr_upload(rp, cols, vals, B)
            s_ref = ctx.scale_l2()
            A_ref, B_ref = ctx.matrix(rmx.F64)
            bufs = [_Dev(hip, rp), _Dev(hip, vals), _Dev(hip, B), _Dev(hip, np.zeros(1))]
            d_rp, d_A, d_B, d_s = bufs
            assert _noalloc(lib, n, d_rp.p, rmx.F64, d_A.p, d_B.p, d_s.p) == 0
            assert hip.hipDeviceSynchronize() == 0
            assert d_s.get()[0] == s_ref
            np.testing.assert_array_equal(d_A.get(), A_ref)
            np.testing.assert_array_equal(d_B.get(), B_ref)
            np.testing.assert_allclose(s_ref, np.linalg.norm(B), rtol=1e-14)
            # FP32 values: A32 / (float) s, the division of CRSMAT::ScaleL2Norm in REALPR float
            A32 = vals.astype(np.float32)
            d_A32, d_B2 = _Dev(hip, A32), _Dev(hip, B)
            assert _noalloc(lib, n, d_rp.p, rmx.F32, d_A32.p, d_B2.p, d_s.p) == 0
            assert hip.hipDeviceSynchronize() == 0
            assert d_s.get()[0] == s_ref
            np.testing.assert_array_equal(d_A32.get(), A32 / np.float32(s_ref))
            np.testing.assert_array_equal(d_B2.get(), B_ref)
            for b in bufs + [d_A32, d_B2]:
                b.free()
        # arguments outside what it takes are refused before anything is queued
        d = _Dev(hip, np.zeros(4))
        assert _noalloc(lib, 0, d.p, rmx.F64, d.p, d.p, d.p) == -1
        assert _noalloc(lib, 4, d.p, 16, d.p, d.p, d.p) == -1
        assert _noalloc(lib, 4, None, rmx.F64, d.p, d.p, d.p) == -1
        d.free()
    finally:
        ctx.close()
